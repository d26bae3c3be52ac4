/*
 * kma_lqueue.h - device side of the unbounded intrusive MPMC queue, HIP for
 * gfx950.  Header-only.
 *
 * The reference's clQueue contract (clQueue.h:42-54, clQueue.cl:40-118):
 * an item's first 8 bytes are its link (clqueue_item.next); kma_lq_enqueue
 * links any number of items (returns 1; 0 only for NULL, an item address
 * above 2^48, or a queue found corrupt), kma_lq_dequeue returns the oldest
 * item or NULL when the queue is empty.  The caller owns an item again once
 * it is dequeued; each lane may name its own queue.
 * (The bounded ring of kma_queue.h is the other pointer queue.)
 *
 * Design: a Michael-Scott queue whose nodes are the items themselves plus
 * one stub node owned by the queue.  head, tail and every link are 64-bit
 * words {tag:16 | pointer:48}; every CAS stores tag + 1, so a word that
 * returns to an old pointer (an item dequeued and enqueued again) no longer
 * matches a stale expected value (ABA).  head names the current dummy node,
 * and a dequeue that swings head from node d to its successor hands d to
 * the caller: items come out in the order they went in, each when its
 * successor has been linked.  When the dummy handed out is the stub, the
 * stub is enqueued again and the dequeue goes on (so the stub never reaches
 * a caller and every item eventually gets a successor).  While a dequeue is
 * putting the stub back, a second dequeue can see an item as the last node
 * and report the queue empty: the reference's dequeue, which also returns
 * NULL without waiting, is no stronger.
 *
 * The lanes of a wavefront that call together (a lane with nothing to
 * enqueue may pass NULL) and name the same queue as the first of them are
 * served together: an enqueue chains their items in lane order and appends
 * the chain with one CAS; a dequeue walks as many links from head as there
 * are callers and takes them all with one CAS (100,000 lanes on one queue
 * then contend as 1,600 wavefronts, not as 100,000 lanes).  Lanes with
 * other queues (each graph node keeps its own) are served one by one, in
 * parallel.  A dequeue still reads the
 * links it takes one after another while head must stay put, so a queue
 * that the whole chip dequeues from at once serves about one item per
 * memory round trip; the ring (kma_queue.h) takes a ticket per lane
 * instead and is the queue for that.  The reference locks a node by
 * swapping POISON into its link (clQueue.cl:97-115) and spins on the tail
 * (clQueue.cl:67); here no node is locked and every retry follows a CAS
 * that another caller won.  A retry loop gives up after 2 s of the wall
 * clock (the queue is then corrupt, or its callers wait on each other far
 * beyond any real use) and the call fails.
 *
 * Visibility: enqueue writes back this XCD's L2 once per wavefront before
 * linking (stores the caller made to the item are visible to a dequeuer on
 * any XCD, the role of the reference's mem_fence calls); dequeue
 * invalidates it once per wavefront after taking items.
 */
#ifndef KMA_LQUEUE_H
#define KMA_LQUEUE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define KMA_LQUEUE_MAGIC 0x3151444c414d4b4bull /* "KKMALDQ1" */
#define KMA_LQUEUE_BYTES 512u

typedef struct kma_dlqueue {
	uint64_t magic;
	uint64_t pad0[15];
	uint64_t head; /* {tag, dummy node} (own 128-B line) */
	uint64_t pad1[15];
	uint64_t tail; /* {tag, last node} (own 128-B line) */
	uint64_t pad2[15];
	uint64_t stub; /* the queue's own node: its link word (own 128-B line) */
	uint64_t pad3[15];
} kma_dlqueue;

namespace kma_lqn {

constexpr uint64_t PTR = (1ull << 48) - 1ull;
__device__ __forceinline__ uint64_t *node(uint64_t w) { return (uint64_t *)(uintptr_t)(w & PTR); }
__device__ __forceinline__ uint64_t tagged(uint64_t ptr, uint64_t old) { return (ptr & PTR) | ((old >> 48) + 1ull) << 48; }
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
struct lqwait {
	uint64_t t0 = 0;
	uint32_t n = 0;
	__device__ __forceinline__ bool expired()
	{
		if ((++n & 63u) != 0)
			return false;
		uint64_t t = wall_clock64();
		if (t0 == 0)
			t0 = t;
		return t - t0 > 200000000ull; /* 2 s */
	}
};

/* Move the tail along links until it names a node with no successor. */
__device__ inline void tail_to_end(uint64_t *tailw, lqwait &wt)
{
	while (!wt.expired()) {
		const uint64_t t = ld(tailw);
		const uint64_t n = ld(node(t)) & PTR;
		if (t != ld(tailw))
			continue;
		if (!n)
			return;
		cas(tailw, t, tagged(n, t));
	}
}

/* Link the chain first .. last (last's link already reset, the chain's
 * own links set) after the last node of the queue whose tail word is given:
 * one CAS on the last node's link, then the tail swung to `last` (a caller
 * that finds the tail behind helps it along, node by node).  A helper can
 * move the tail onto `first` before the swing; the tail is then walked along
 * the chain so that it never stays behind a call that has returned. */
__device__ inline bool link_last(uint64_t *tailw, uint64_t *first, uint64_t *last)
{
	lqwait wt;
	for (;;) {
		if (wt.expired())
			return false;
		const uint64_t t = ld(tailw);
		uint64_t *end = node(t);
		const uint64_t n = ld(end);
		if (t != ld(tailw))
			continue;
		if ((n & PTR) == 0) {
			if (cas(end, n, tagged((uint64_t)(uintptr_t)first, n))) {
				if (!cas(tailw, t, tagged((uint64_t)(uintptr_t)last, t)) && first != last)
					tail_to_end(tailw, wt);
				return true;
			}
		} else {
			cas(tailw, t, tagged(n, t)); /* help a linked node become the tail */
		}
	}
}

__device__ __forceinline__ void reset_link(uint64_t *it)
{
	st(it, tagged(0, ld(it)));
}

/* lanes strictly below l (l < 64) */
__device__ __forceinline__ uint64_t lanes_below(uint32_t l) { return (1ull << l) - 1ull; }

__device__ __forceinline__ bool valid_item(const void *item)
{
	return item != nullptr && ((uint64_t)(uintptr_t)item & ~PTR) == 0;
}

/* The lanes whose queue is the first valid calling lane's are served
 * together: their items are chained in lane order with plain stores (no
 * contention), written back once, and that first lane appends the chain
 * with one CAS.  Lanes with another queue (every node of a graph has its
 * own) link their item one by one, in parallel. */
__device__ inline int enqueue_at(uint64_t *tailw, void *item)
{
	const uint32_t me = __lane_id();
	const bool ok = valid_item(item);
	const uint64_t valid = __ballot(ok);
	if (!valid)
		return 0;
	const uint32_t lead = (uint32_t)__ffsll((unsigned long long)valid) - 1u;
	const uint64_t qt = __shfl((uint64_t)(uintptr_t)tailw, (int)lead);
	const bool in = ok && (uint64_t)(uintptr_t)tailw == qt;
	const uint64_t mask = __ballot(in);
	const uint64_t above = mask & ~lanes_below(me) & ~(1ull << me);
	const uint32_t nl = above ? (uint32_t)__ffsll((unsigned long long)above) - 1u : me;
	const uint64_t nxt = __shfl((uint64_t)(uintptr_t)item, (int)nl);
	uint64_t *it = (uint64_t *)item;
	if (ok)
		st(it, tagged(in && above ? nxt : 0u, ld(it)));
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
	const uint32_t lastl = 63u - (uint32_t)__clzll((unsigned long long)mask);
	const uint64_t lastp = __shfl((uint64_t)(uintptr_t)item, (int)lastl);
	int r = 0;
	if (me == lead)
		r = link_last(tailw, it, (uint64_t *)(uintptr_t)lastp) ? 1 : 0;
	else if (ok && !in)
		r = link_last(tailw, it, it) ? 1 : 0;
	const int rl = __shfl(r, (int)lead);
	return in ? rl : r;
}

/* One item for one lane (the lanes whose queue differs from the served
 * group's): the plain Michael-Scott dequeue, stub put back when it comes
 * round. */
__device__ inline void *dequeue_one(uint64_t *headw, uint64_t *tailw, uint64_t *stub)
{
	lqwait wt;
	while (!wt.expired()) {
		const uint64_t h = ld(headw);
		const uint64_t t = ld(tailw);
		const uint64_t nx = ld(node(h)) & PTR;
		if (h != ld(headw))
			continue;
		if (!nx)
			return nullptr; /* only the dummy: empty */
		if ((h & PTR) == (t & PTR)) {
			cas(tailw, t, tagged(nx, t)); /* the tail must not stay on the node taken */
			continue;
		}
		if (!cas(headw, h, tagged(nx, h)))
			continue;
		if ((h & PTR) == (uint64_t)(uintptr_t)stub) {
			reset_link(stub);
			if (!link_last(tailw, stub, stub))
				return nullptr;
			continue;
		}
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
		return node(h);
	}
	return nullptr;
}

/* The lanes whose queue is the first calling lane's ask together: the first asking lane reads head and walks as many links as
 * there are askers, broadcasting the j-th node to the j-th asker; the tail
 * is helped past the nodes taken; then ONE CAS moves head past them all.  A
 * lane handed the stub puts it back and asks again; lanes left over when the
 * chain ran short get NULL (empty).  Lanes with another queue dequeue one
 * by one, in parallel. */
__device__ inline void *dequeue_at(uint64_t *headw, uint64_t *tailw, uint64_t *stub)
{
	const uint32_t me = __lane_id();
	const uint32_t l0 = (uint32_t)__ffsll((unsigned long long)__ballot(1)) - 1u;
	const uint64_t q0 = __shfl((uint64_t)(uintptr_t)headw, (int)l0);
	const bool in = (uint64_t)(uintptr_t)headw == q0;
	uint64_t want = __ballot(in);
	void *ret = nullptr;
	lqwait wt;
	while (want) {
		if (wt.expired())
			break;
		const uint32_t lead = (uint32_t)__ffsll((unsigned long long)want) - 1u;
		const uint32_t m = (uint32_t)__popcll(want);
		const bool asking = (want >> me) & 1ull;
		const uint32_t rk = (uint32_t)__popcll(want & lanes_below(me));
		uint64_t h = 0, t = 0;
		if (me == lead) {
			h = ld(headw);
			t = ld(tailw);
		}
		h = __shfl(h, (int)lead);
		t = __shfl(t, (int)lead);
		/* the walk: a link is followed only while head still holds h (same
		 * tag), i.e. while no dequeue has taken the node it was read from,
		 * so every node reached is still in the queue (its first word is a
		 * link, not a caller's data) */
		uint64_t cur = h & PTR, mine = 0;
		uint32_t k = 0;
		int stale = 0;
		for (uint32_t j = 0; j < m; j++) {
			uint64_t nx = 0;
			if (me == lead) {
				nx = ld((const uint64_t *)(uintptr_t)cur) & PTR;
				stale = ld(headw) != h;
			}
			nx = __shfl(nx, (int)lead);
			if (__shfl(stale, (int)lead))
				break;
			if (!nx)
				break;
			if (asking && rk == j)
				mine = cur;
			cur = nx;
			k = j + 1u;
		}
		if (__shfl(stale, (int)lead))
			continue;
		if (k == 0)
			break; /* only the dummy: empty */
		/* the tail must not be left on a node taken: help it one node on
		 * (to its own successor) and look again */
		if (__ballot(mine != 0 && mine == (t & PTR))) {
			if (me == lead) {
				const uint64_t tn = ld(node(t)) & PTR;
				if (tn)
					cas(tailw, t, tagged(tn, t));
			}
			continue;
		}
		int won = 0;
		if (me == lead)
			won = cas(headw, h, tagged(cur, h)) ? 1 : 0;
		if (!__shfl(won, (int)lead))
			continue;
		if (mine == (uint64_t)(uintptr_t)stub) {
			/* the stub went round: put it back behind the last node */
			reset_link(stub);
			if (!link_last(tailw, stub, stub))
				mine = ~0ull; /* corrupt queue: give up below */
			else
				mine = 0;
		}
		if (mine && mine != ~0ull)
			ret = (void *)(uintptr_t)mine;
		/* askers served, or beyond a short chain (empty for them) */
		const uint64_t again = __ballot(asking && mine == 0 && rk < k);
		want = again;
	}
	if (__ballot(ret != nullptr))
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
	if (!in)
		ret = dequeue_one(headw, tailw, stub);
	return ret;
}

} // namespace kma_lqn

/* Append item (its first 8 bytes become its link); 1 iff it was linked. */
__device__ inline int kma_lq_enqueue(kma_dlqueue *q, void *item)
{
	return kma_lqn::enqueue_at(&q->tail, item);
}

/* The oldest item, or NULL if the queue is empty. */
__device__ inline void *kma_lq_dequeue(kma_dlqueue *q)
{
	return kma_lqn::dequeue_at(&q->head, &q->tail, &q->stub);
}

/* The same queue in 24 bytes, to embed in a record (the reference keeps a
 * clqueue in every graph node, tb_clTree.cl:67): head, tail and the stub's
 * link word, with no padding.  kma_lq_init must run before any other call
 * can see the record. */
typedef struct kma_lq {
	uint64_t head, tail, stub;
} kma_lq;

__device__ __forceinline__ void kma_lq_init(kma_lq *q)
{
	using namespace kma_lqn;
	st(&q->stub, 0);
	st(&q->head, (uint64_t)(uintptr_t)&q->stub);
	st(&q->tail, (uint64_t)(uintptr_t)&q->stub);
}
__device__ inline int kma_lq_push(kma_lq *q, void *item)
{
	return kma_lqn::enqueue_at(&q->tail, item);
}
__device__ inline void *kma_lq_pop(kma_lq *q)
{
	return kma_lqn::dequeue_at(&q->head, &q->tail, &q->stub);
}
/* Pointer part of a head, tail or link word. */
__device__ __forceinline__ uint64_t kma_lq_ptr(uint64_t w) { return w & kma_lqn::PTR; }

#endif /* KMA_LQUEUE_H */
