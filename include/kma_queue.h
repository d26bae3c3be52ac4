/*
 * kma_queue.h - device side of the MPMC pointer queue, HIP for gfx950.
 * Header-only.
 *
 * Same call surface as the reference's clQueue (clQueue.cl:40-118):
 * kma_enqueue(q, item) returns 1 iff the item was added (0 for NULL or a
 * full queue); kma_dequeue(q) returns the oldest item or NULL when the queue
 * is empty.  Any lane may call either, with no uniformity requirement.
 *
 * Design: a bounded ring of `cap` slots {seq, ptr} with 64-bit head and tail
 * tickets (Vyukov's sequence-number ring).  Slot i of lap L is free for the
 * enqueue of ticket t = L*cap + i when seq == t and holds that ticket's item
 * when seq == t + 1; the dequeue of t sets seq = t + cap.  The reference
 * links items through their first word and locks a node with a POISON bit
 * (clQueue.cl:97-115); the ring needs no per-item field, no ABA tag and no
 * lock, and the lanes of a wavefront that call together move the tail or
 * head with ONE compare-and-swap for all of them: each lane checks its own
 * slot (lane k -> ticket pos + k), the leader claims the longest run of
 * ready slots, and lanes past it retry.
 *
 * A full ring makes kma_enqueue return 0; so can a ring with one free slot
 * whose dequeue is still completing, so size the ring above the most items
 * it must hold.
 *
 * Visibility: enqueue writes back this XCD's L2 once per wavefront before it
 * publishes (so stores the caller made to the item are visible to a
 * dequeuer on any XCD, the role of the reference's mem_fence calls), and
 * dequeue invalidates this XCD's L2 once per wavefront after it took items.
 */
#ifndef KMA_QUEUE_H
#define KMA_QUEUE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define KMA_QUEUE_MAGIC 0x3151414d4b4d414bull /* "KAMKMAQ1" */
#define KMA_QUEUE_HDR 384u                    /* slots start here */

typedef struct kma_dqueue_slot {
	uint64_t seq;
	uint64_t ptr;
} kma_dqueue_slot;

typedef struct kma_dqueue {
	uint64_t magic;
	uint64_t cap; /* power of two */
	uint64_t pad0[14];
	uint64_t head; /* next ticket to dequeue (own 128-B line) */
	uint64_t pad1[15];
	uint64_t tail; /* next ticket to enqueue (own 128-B line) */
	uint64_t pad2[15];
} kma_dqueue;

namespace kma_q {

__device__ __forceinline__ kma_dqueue_slot *slot(kma_dqueue *q, uint64_t t)
{
	return (kma_dqueue_slot *)((char *)q + KMA_QUEUE_HDR) + (t & (q->cap - 1u));
}
__device__ __forceinline__ uint64_t ld(const uint64_t *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st(uint64_t *p, uint64_t v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas(uint64_t *p, uint64_t cmp, uint64_t val)
{
	return __hip_atomic_compare_exchange_strong(p, &cmp, val, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
						    __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void wait_prior()
{
	__atomic_signal_fence(__ATOMIC_SEQ_CST);
	__builtin_amdgcn_s_waitcnt(0);
	__atomic_signal_fence(__ATOMIC_SEQ_CST);
}
/* lanes of `grp` strictly below lane l */
__device__ __forceinline__ uint32_t rank_in(uint64_t grp, uint32_t l)
{
	return (uint32_t)__popcll(grp & ((1ull << l) - 1ull));
}
/* Retry loops give up after 0.25 s of the 100 MHz wall clock (a corrupt
 * queue): the calls then fail instead of spinning on. */
struct qwait {
	uint64_t t0 = 0;
	uint32_t n = 0;
	__device__ __forceinline__ bool expired()
	{
		if ((++n & 63u) != 0)
			return false;
		uint64_t t = wall_clock64();
		if (t0 == 0)
			t0 = t;
		return t - t0 > 25000000ull;
	}
};
/* number of leading members of `grp` (in lane order) whose bit is set in
 * `ok` */
__device__ __forceinline__ uint32_t ready_run(uint64_t grp, uint64_t ok)
{
	uint64_t bad = grp & ~ok;
	if (!bad)
		return (uint32_t)__popcll(grp);
	return rank_in(grp, (uint32_t)__ffsll((unsigned long long)bad) - 1u);
}

} // namespace kma_q

__device__ inline int kma_enqueue(kma_dqueue *q, void *item)
{
	using namespace kma_q;
	const uint32_t me = __lane_id();
	bool pending = item != nullptr;
	int ret = 0;
	if (__ballot(pending))
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
	qwait wt;
	for (;;) {
		uint64_t grp = __ballot(pending);
		if (!grp || wt.expired())
			break;
		uint32_t leader = (uint32_t)__ffsll((unsigned long long)grp) - 1u;
		uint64_t pos = 0;
		if (me == leader)
			pos = ld(&q->tail);
		pos = __shfl(pos, (int)leader);
		uint32_t r = rank_in(grp, me);
		uint64_t seq = pending ? ld(&slot(q, pos + r)->seq) : 0;
		uint64_t ok = __ballot(pending && seq == pos + r);
		uint32_t k = ready_run(grp, ok);
		if (k == 0) {
			/* the leader's slot still holds an item of the previous lap:
			 * full; otherwise the tail moved on, retry */
			uint64_t s0 = __shfl(seq, (int)leader);
			if ((int64_t)(s0 - pos) < 0)
				break;
			continue;
		}
		int won = 0;
		if (me == leader)
			won = cas(&q->tail, pos, pos + k) ? 1 : 0;
		won = __shfl(won, (int)leader);
		if (won && pending && r < k) {
			kma_dqueue_slot *s = slot(q, pos + r);
			st(&s->ptr, (uint64_t)(uintptr_t)item);
			wait_prior();
			st(&s->seq, pos + r + 1u);
			pending = false;
			ret = 1;
		}
	}
	return ret;
}

__device__ inline void *kma_dequeue(kma_dqueue *q)
{
	using namespace kma_q;
	const uint32_t me = __lane_id();
	bool pending = true;
	void *out = nullptr;
	bool took = false;
	qwait wt;
	for (;;) {
		uint64_t grp = __ballot(pending);
		if (!grp || wt.expired())
			break;
		uint32_t leader = (uint32_t)__ffsll((unsigned long long)grp) - 1u;
		uint64_t pos = 0;
		if (me == leader)
			pos = ld(&q->head);
		pos = __shfl(pos, (int)leader);
		uint32_t r = rank_in(grp, me);
		uint64_t seq = pending ? ld(&slot(q, pos + r)->seq) : 0;
		uint64_t ok = __ballot(pending && seq == pos + r + 1u);
		uint32_t k = ready_run(grp, ok);
		if (k == 0) {
			/* the leader's slot is not filled for this lap: empty (or an
			 * enqueue of that ticket is still writing it) */
			uint64_t s0 = __shfl(seq, (int)leader);
			if ((int64_t)(s0 - (pos + 1u)) < 0)
				break;
			continue;
		}
		int won = 0;
		if (me == leader)
			won = cas(&q->head, pos, pos + k) ? 1 : 0;
		won = __shfl(won, (int)leader);
		if (won && pending && r < k) {
			kma_dqueue_slot *s = slot(q, pos + r);
			out = (void *)(uintptr_t)ld(&s->ptr);
			wait_prior();
			st(&s->seq, pos + r + q->cap);
			pending = false;
			took = true;
		}
	}
	if (__ballot(took))
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
	return out;
}

#endif /* KMA_QUEUE_H */
