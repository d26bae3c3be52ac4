/*
 * kma_wg_device.h - work-group collective malloc/free on the native layout
 * (HIP, gfx950).  Header-only.
 *
 * kma_wg_malloc / kma_wg_free are called by EVERY thread of a 1-D work-group
 * (uniform control flow, like the reference's WG-collective
 * clArrayList_grow_local, clArrayList.cl:111-176).  Each thread passes its own
 * request (size 0 / NULL: none) and gets its own result.  The aggregation
 * width AGG (a template argument) says over how many requests the
 * device-wide atomics are combined:
 *   KMA_AGG_WAVE (default): each wavefront serves its own lanes with the
 *           wavefront algorithm of kma_malloc / kma_free (lanes grouped by
 *           class with ballots, one reservation per class group, one bitmap
 *           RMW per word of consecutive reserved slots); no LDS, no barrier,
 *           so the wavefronts of a work-group never wait for each other;
 *   G = 8, 16 or 32: the work-group stages the whole batch in LDS:
 *     malloc: LDS class histogram -> one reservation (atomicAdd on the
 *             superblock state word) per class; a G-lane group per class
 *             (all classes of the batch at once) reads the superblock bitmap
 *             coalesced (lane i <- word i), prefix-sums the free bits and
 *             claims them with one atomicOr per word;
 *     free:   an LDS hash keyed by (superblock, bitmap word) ORs the bits of
 *             all requests that share a word; one atomicAnd per word and one
 *             atomicSub on the state word per entry.
 *   Work-group staging combines more requests per atomic (a class with a
 *   superblock's worth of requests takes whole superblocks) at the price of
 *   barriers; which width wins depends on the batch, and the stream
 *   autotuner picks it.
 * Requests larger than the largest class go to the page allocator per lane.
 */
#ifndef KMA_WG_DEVICE_H
#define KMA_WG_DEVICE_H

#include "kma_device.h"

template <int BLOCK>
struct kma_wg_lds {
	uint32_t cnt[32];
	uint32_t cbase[33];
	uint16_t order[BLOCK];
	uint64_t res[BLOCK];
	unsigned long long hkey[2 * BLOCK];
	uint32_t hmask[2 * BLOCK];
	uint32_t hn[2 * BLOCK];
	/* per class: the context head value this work-group's last reservation
	 * of the class came from (0: none); the next reservation tries that
	 * superblock before reading the head.  kma_wg_lds_init clears it;
	 * left uninitialised it only costs first attempts (any value is
	 * checked like a stale head read). */
	uint64_t scache[32];
};

/* Clear the work-group's superblock cache; every thread calls it once,
 * before the work-group's first collective call. */
template <int BLOCK>
__device__ __forceinline__ void kma_wg_lds_init(kma_wg_lds<BLOCK> &L)
{
	if (threadIdx.x < 32u)
		L.scache[threadIdx.x] = 0;
	__syncthreads();
}

namespace kma_nat {

/* The lowest `k` set bits of `m`. */
__device__ __forceinline__ uint32_t low_bits(uint32_t m, uint32_t k)
{
	uint32_t r = 0;
	for (uint32_t i = 0; i < k; i++) {
		uint32_t b = m & (0u - m);
		r |= b;
		m ^= b;
	}
	return r;
}

/* ---- G-lane groups: one size class per group, 64/G per wavefront ----
 * Every lane of the wavefront executes these (the loops are
 * wavefront-uniform); a group with nothing to do passes dummies. */
template <uint32_t G>
__device__ __forceinline__ uint32_t grp_lane() { return lane() & (G - 1u); }

template <uint32_t G>
__device__ __forceinline__ uint32_t group_excl_scan(uint32_t v, uint32_t *total)
{
	uint32_t x = v, gl = grp_lane<G>();
	for (uint32_t o = 1; o < G; o <<= 1) {
		uint32_t y = __shfl_up(x, o, (int)G);
		if (gl >= o)
			x += y;
	}
	*total = __shfl(x, (int)G - 1, (int)G);
	return x - v;
}

template <uint32_t G>
__device__ __forceinline__ uint32_t group_bcast(uint32_t v)
{
	return __shfl(v, 0, (int)G);
}

/* Each group with want set claims `got` free bits of its superblock sb
 * (class c) and writes block offsets (from the heap base) into
 * res[order[0..done-1]]; returns the group's `done`.  The first pass claims
 * the reserved slots [base, base + got) themselves with no read (blocks are
 * claimed bottom-up and mostly freed in bulk, so they are usually free; the
 * RMWs say which were).  Later passes read the words from the one holding
 * bit `base`, 16 words per pass, and take the lowest free bits; after a full
 * pass over the bitmap without progress the words are re-read with RMWs. */
template <uint32_t G>
__device__ inline uint32_t group_claim(const kma_dheap *H, bool want, uint32_t sb, uint32_t c, uint32_t base,
				       uint32_t got, const uint16_t *order, uint64_t *res)
{
	const uint32_t gl = grp_lane<G>();
	uint32_t cap = want ? H->class_cap[c] : 32u;
	uint32_t W = (cap + 31u) >> 5;
	uint32_t tail = (cap & 31u) ? ((1u << (cap & 31u)) - 1u) : ~0u;
	uint32_t *bm = want ? bitmap(H, sb) : nullptr;
	uint64_t d0 = want ? H->off_data + ((uint64_t)sb << H->sb_log2) : 0u;
	uint32_t bsize = want ? H->class_size[c] : 0u;
	uint32_t w0 = want ? (base >> 5) % W : 0u;
	uint32_t done = 0, chunk = 0, before = 0;
	bool fresh = false, opt = true, need = want && got > 0;
	waiter wt;
	while (__ballot(need)) {
		uint32_t k = chunk + gl;
		bool has = need && k < W;
		uint32_t wi = has ? (w0 + k) % W : 0u;
		uint32_t freeb = 0;
		if (has) {
			if (opt)
				freeb = range_bits(wi, base, min(base + got, cap));
			else
				freeb = ~(fresh ? f_or(bm + wi, 0u) : ld(bm + wi));
			if (wi == W - 1u)
				freeb &= tail;
		}
		uint32_t tot;
		uint32_t pre = group_excl_scan<G>(__popc(freeb), &tot);
		uint32_t take = need ? min((uint32_t)__popc(freeb), sat_sub(got - done, pre)) : 0u;
		uint32_t mask = low_bits(freeb, take);
		uint32_t won = 0;
		if (mask)
			won = mask & ~f_or(bm + wi, mask);
		uint32_t wtot;
		uint32_t wpre = group_excl_scan<G>(__popc(won), &wtot);
		uint32_t r = done + wpre;
		while (won) {
			uint32_t b = __ffs(won) - 1u;
			won &= won - 1u;
			res[order[r++]] = d0 + (uint64_t)(wi * 32u + b) * bsize;
		}
		const bool exp = group_bcast<G>(need && wt.expired() ? 1u : 0u) != 0;
		if (need) {
			done += wtot;
			if (opt)
				opt = false; /* then scan from word w0 */
			else
				chunk += G;
			if (chunk >= W) {
				/* a full pass: re-read with RMWs if it made no progress */
				chunk = 0;
				fresh = done == before;
				before = done;
			}
			if (done >= got || exp) {
				if (done < got && gl == 0)
					bump_stat(H, KMA_STAT_WAIT_WGCLAIM);
				need = false;
			}
		}
	}
	return done;
}

/* Whole-superblock grab (every lane of the wavefront calls; lanes with
 * want set each pop one superblock for class c and claim all `cap` slots).
 * The successful lanes of each group take consecutive rank ranges of `cap`
 * from rank 0; returns the group's number of ranks covered.  A pooled
 * superblock has count 0 and a clear bitmap, so every RMW should win; a
 * slot that does not (a corrupt heap) is left out and its count given back,
 * and its rank gets NULL. */
#define KMA_WHOLE_WORDS 8u
template <uint32_t G>
__device__ __forceinline__ uint32_t whole_superblocks(const kma_dheap *H, uint32_t x, uint32_t ctx, uint32_t c,
					     uint32_t cap, bool want, const uint16_t *order, uint64_t *res)
{
	uint32_t sb = 0;
	bool have = want && pool_pop(H, x, ctx, grp_lane<G>(), &sb);
	uint32_t old[KMA_WHOLE_WORDS];
	const uint32_t W = (cap + 31u) >> 5;
	const uint32_t tail = (cap & 31u) ? ((1u << (cap & 31u)) - 1u) : ~0u;
	if (have) {
		fadd(state(H, sb), ((uint64_t)(c + 1u) << ST_CLASS_SHIFT) | ((uint64_t)x << ST_HOME_SHIFT) | cap);
		uint32_t *bm = bitmap(H, sb);
#pragma unroll
		for (uint32_t i = 0; i < KMA_WHOLE_WORDS; i++)
			if (i < W)
				old[i] = f_or(bm + i, i == W - 1u ? tail : ~0u);
	}
	uint32_t tot;
	const uint32_t idx = group_excl_scan<G>(have ? 1u : 0u, &tot);
	if (have) {
		const uint64_t d0 = H->off_data + ((uint64_t)sb << H->sb_log2);
		const uint32_t bsize = H->class_size[c];
		uint32_t lost = 0;
#pragma unroll
		for (uint32_t i = 0; i < KMA_WHOLE_WORDS; i++) {
			if (i >= W)
				continue;
			uint32_t full = i == W - 1u ? tail : ~0u;
			for (uint32_t b = 0; b < 32u && i * 32u + b < cap; b++) {
				bool won = ((full & ~old[i]) >> b) & 1u;
				lost += won ? 0u : 1u;
				res[order[idx * cap + i * 32u + b]] = won ? d0 + (uint64_t)(i * 32u + b) * bsize : 0u;
			}
		}
		if (lost) {
			bump_stat(H, KMA_STAT_WAIT_WGCLAIM);
			undo(H, sb, lost, x, ctx);
		}
	}
	return tot * cap;
}

/* Partial grab (every lane of the wavefront calls; lanes of a group with
 * need > 0 take part): under pool pressure a class still short of `need`
 * blocks takes listed superblocks straight from the partial index, one per
 * group lane (as many lanes as need / an eighth of a superblock, all in
 * flight at once; taking more only returns them, and while they are out
 * other callers find the index emptier).  The group prefix-sums their free slots and each lane
 * takes only what the group still needs: one atomicAdd on the state word
 * for the slots, then the lowest free bits of its bitmap words with one
 * atomicOr per word.  Results go to res[order[0..done-1]] in lane order;
 * returns the group's `done`.  Each superblock then goes back into the
 * index or out of it (relist). */
template <uint32_t G>
__device__ inline uint32_t partial_superblocks(const kma_dheap *H, uint32_t x, uint32_t ctx, uint32_t c, uint32_t need,
					       const uint16_t *order, uint64_t *res)
{
	const uint32_t gl = grp_lane<G>();
	const uint32_t cap = need ? H->class_cap[c] : 32u, room = part_room(cap);
	const uint32_t lanes = need ? min(G, (need + room - 1u) / room) : 0u;
	uint32_t sb = 0;
	const bool have = gl < lanes && pidx_take(H, x, c, ctx * 37u + gl * 11u, &sb);
	const uint32_t fr = have ? sat_sub(cap, st_count(ld(state(H, sb)))) : 0u;
	uint32_t tot;
	const uint32_t pre = group_excl_scan<G>(fr, &tot);
	const uint32_t take = have ? min(fr, sat_sub(need, pre)) : 0u;
	uint32_t got = 0;
	if (take) {
		const uint64_t old = fadd(state(H, sb), (uint64_t)take);
		got = min(take, sat_sub(cap, st_count(old)));
		if (got < take)
			undo(H, sb, take - got, x, ctx);
		bump_stat(H, KMA_STAT_REATTACHED);
	}
	/* the lowest free bits of the superblock's words, up to `got` */
	const uint32_t W = (cap + 31u) >> 5;
	const uint32_t tail = (cap & 31u) ? ((1u << (cap & 31u)) - 1u) : ~0u;
	uint32_t won[KMA_WHOLE_WORDS];
	uint32_t nwon = 0;
	uint32_t *bm = got ? bitmap(H, sb) : nullptr;
#pragma unroll
	for (uint32_t i = 0; i < KMA_WHOLE_WORDS; i++) {
		won[i] = 0;
		if (got > nwon && i < W) {
			const uint32_t fb = ~ld(bm + i) & (i == W - 1u ? tail : ~0u);
			const uint32_t pick = low_bits(fb, got - nwon);
			if (pick) {
				won[i] = pick & ~f_or(bm + i, pick);
				nwon += __popc(won[i]);
			}
		}
	}
	if (got > nwon)
		undo(H, sb, got - nwon, x, ctx); /* raced by a reservation still in flight */
	uint32_t wtot;
	uint32_t r = group_excl_scan<G>(nwon, &wtot);
	if (nwon) {
		const uint64_t d0 = H->off_data + ((uint64_t)sb << H->sb_log2);
		const uint32_t bsize = H->class_size[c];
#pragma unroll
		for (uint32_t i = 0; i < KMA_WHOLE_WORDS; i++)
			for (uint32_t m = won[i]; m; m &= m - 1u)
				res[order[r++]] = d0 + (uint64_t)(i * 32u + (uint32_t)(__ffs(m) - 1)) * bsize;
	}
	if (have)
		relist(H, sb, x, ctx, true);
	return wtot;
}
/* Rounds of partial grabs per class and call under pool pressure; the
 * head refills (which also take listed superblocks) serve what is left.
 * One round measured 3-7 % faster than four on the oversubscribed request
 * stream, and the serving loop completed 3 % more jobs with a lower p99. */
#ifndef KMA_PART_ROUNDS
#define KMA_PART_ROUNDS 1u
#endif

__device__ __forceinline__ uint32_t wg_hash(uint64_t key, uint32_t n)
{
	return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 40) & (n - 1u);
}

} /* namespace kma_nat */

/* Aggregation widths (see the top of this file). */
#define KMA_AGG_WAVE 0u

/* Work-group collective malloc.  All threads of the 1-D work-group call it;
 * size 0 means "no request" for that thread.  With work-group aggregation,
 * per_lane_claim selects one bitmap atomic per block instead of one per
 * bitmap word. */
template <int BLOCK, uint32_t AGG = KMA_AGG_WAVE>
__device__ __forceinline__ void *kma_wg_malloc(kma_dheap *H, size_t size, kma_wg_lds<BLOCK> &L, uint32_t per_lane_claim = 0)
{
	using namespace kma_nat;
	static_assert(AGG == KMA_AGG_WAVE || AGG == 8u || AGG == 16u || AGG == 32u, "aggregation width");
	if constexpr (AGG == KMA_AGG_WAVE) {
		(void)L;
		(void)per_lane_claim;
		return kma_malloc(H, size);
	}
	const uint32_t tid = threadIdx.x, nth = blockDim.x, nw = (nth + 63u) >> 6;
	const uint32_t x = xcc(), ctx = wg_ctx_of(H, x);
	const uint32_t ncls = H->nclasses;
	const bool press = under_pressure(H, x);
	if (tid < 32)
		L.cnt[tid] = 0;
	__syncthreads();
	int c = -1;
	if (size > 0 && size <= H->max_small)
		c = class_of(H, (uint32_t)size);
	uint32_t q = 0;
	if (c >= 0)
		q = atomicAdd(&L.cnt[c], 1u);
	__syncthreads();
	if (tid == 0) {
		uint32_t acc = 0;
		for (uint32_t k = 0; k < ncls; k++) {
			L.cbase[k] = acc;
			acc += L.cnt[k];
		}
		L.cbase[ncls] = acc;
	}
	__syncthreads();
	if (c >= 0)
		L.order[L.cbase[c] + q] = (uint16_t)tid;
	__syncthreads();
	/* one 16-lane group per class, all classes at once: the group's first
	 * lane reserves a segment, the group claims it; repeat until the class is
	 * served or the heap cannot serve it (loops are wavefront-uniform) */
	constexpr uint32_t G = AGG ? AGG : 16u;
	const uint32_t ngroups = nw * (64u / G), g = tid / G, gl = tid % G;
	const uint32_t iters = (ncls + ngroups - 1u) / ngroups;
	for (uint32_t it = 0; it < iters; it++) {
		const uint32_t k = g + it * ngroups;
		const uint32_t n = k < ncls ? L.cnt[k] : 0u;
		const uint16_t *ord = n ? &L.order[L.cbase[k]] : L.order;
		uint32_t r0 = 0;
		/* a class with at least a superblock's worth of requests takes whole
		 * superblocks straight from the pool, one per group lane, all in
		 * flight at once, and claims every slot with no reads */
		const uint32_t cap = n ? H->class_cap[k] : 0u;
		const uint32_t nwhole = (n && cap > 1u && cap <= 32u * KMA_WHOLE_WORDS) ? min(n / cap, G) : 0u;
		if (__ballot(nwhole != 0u) && !press)
			r0 = whole_superblocks<G>(H, x, ctx, k, cap, gl < nwhole, ord, L.res);
		/* under pool pressure the class takes listed superblocks instead,
		 * several at once, before any head refill */
		for (uint32_t rd = 0; press && rd < KMA_PART_ROUNDS; rd++) {
			const uint32_t need = r0 < n && cap > 1u && cap <= 32u * KMA_WHOLE_WORDS ? n - r0 : 0u;
			if (!__ballot(need != 0u))
				break;
			const uint32_t done = partial_superblocks<G>(H, x, ctx, k, need, ord + r0, L.res);
			r0 += done;
			if (!__ballot(need != 0u && done != 0u))
				break;
		}
		bool pend = r0 < n;
		waiter wt;
		while (__ballot(pend)) {
			if (__ballot(wt.expired())) {
				/* a head stayed busy for the whole watchdog period */
				if (pend && gl == 0)
					bump_stat(H, KMA_STAT_WAIT_RESERVE);
				break;
			}
			resv r = {0, 0, 0, 0};
			if (pend && gl == 0) {
				uint64_t used;
				r = reserve(H, x, ctx, k, n - r0, L.scache[k], &used);
				L.scache[k] = used;
			}
			const uint32_t ok = group_bcast<G>(r.ok), sb = group_bcast<G>(r.sb);
			const uint32_t rb = group_bcast<G>(r.base), got = group_bcast<G>(r.got);
			const bool want = pend && ok;
			if (pend && !ok)
				pend = false;
			uint32_t done;
			if (per_lane_claim) {
				/* one claim per block: group lane j takes ranks j, j+16, ... */
				uint32_t mine = 0;
				if (want)
					for (uint32_t q = gl; q < got; q += G) {
						char *p = claim(H, sb, k, rb + q, gl);
						L.res[ord[r0 + q]] = p ? (uint64_t)(p - base(H)) : 0;
						mine += p ? 1u : 0u;
					}
				uint32_t tot;
				group_excl_scan<G>(mine, &tot);
				done = tot;
			} else {
				done = group_claim<G>(H, want, sb, k, rb, got, ord + r0, L.res);
			}
			if (want) {
				r0 += done;
				if (done < got) {
					/* give back what could not be claimed */
					if (gl == 0)
						undo(H, sb, got - done, x, ctx);
					pend = false;
				} else if (r0 >= n) {
					pend = false;
				}
			}
		}
		for (uint32_t r = r0 + gl; r < n; r += G)
			L.res[ord[r]] = 0;
	}
	__syncthreads();
	void *out = nullptr;
	if (c >= 0) {
		uint64_t off = L.res[tid];
		out = off ? (void *)(base(H) + off) : nullptr;
	} else if (size > H->max_small) {
		out = page_alloc(H, size, (blockIdx.x * nth + tid) ^ (uint32_t)clock64());
	}
	__syncthreads();
	return out;
}

/* Work-group collective free.  All threads call it; NULL means "no
 * request".  Double frees are detected and counted, not applied. */
template <int BLOCK, uint32_t AGG = KMA_AGG_WAVE>
__device__ __forceinline__ void kma_wg_free(kma_dheap *H, void *ptr, kma_wg_lds<BLOCK> &L)
{
	using namespace kma_nat;
	if constexpr (AGG == KMA_AGG_WAVE) {
		(void)L;
		kma_free(H, ptr);
		return;
	}
	const uint32_t tid = threadIdx.x, nth = blockDim.x;
	const uint32_t x = xcc(), ctx = wg_ctx_of(H, x);
	const uint32_t HN = 2u * BLOCK;
	/* every thread's stores to its block complete before any bit of the
	 * batch is released (the barrier below only orders LDS) */
	order_prior();
	for (uint32_t i = tid; i < HN; i += nth) {
		L.hkey[i] = 0;
		L.hmask[i] = 0;
		L.hn[i] = 0;
	}
	__syncthreads();
	bool owner = false;
	uint32_t slot = 0;
	if (ptr) {
		char *p = (char *)ptr;
		char *d0 = base(H) + H->off_data;
		uint64_t off = (uint64_t)(p - d0);
		if (p >= d0 && off < (H->nsb << H->sb_log2)) {
			uint32_t sb = (uint32_t)(off >> H->sb_log2);
			uint32_t c1 = st_cls1(ld(state(H, sb)));
			if (c1 == 0) {
				bump_stat(H, KMA_STAT_BAD_FREE);
			} else {
				uint64_t inoff = off & ((1ull << H->sb_log2) - 1ull);
				uint32_t idx = (uint32_t)((inoff * H->class_magic[c1 - 1u]) >> 40);
				unsigned long long key = (unsigned long long)sb * H->bm_words + (idx >> 5) + 1ull;
				uint32_t h = wg_hash(key, HN);
				for (;;) {
					unsigned long long old = atomicCAS(&L.hkey[h], 0ull, key);
					if (old == 0ull || old == key) {
						owner = old == 0ull;
						slot = h;
						break;
					}
					h = (h + 1u) & (HN - 1u);
				}
				atomicOr(&L.hmask[slot], 1u << (idx & 31u));
				atomicAdd(&L.hn[slot], 1u);
			}
		} else if (H->npages && p >= base(H) + H->off_pages) {
			page_free(H, p);
		} else {
			bump_stat(H, KMA_STAT_BAD_FREE);
		}
	}
	__syncthreads();
	uint32_t sb = 0, wi = 0, nrem = 0;
	if (owner) {
		unsigned long long key = L.hkey[slot] - 1ull;
		sb = (uint32_t)(key / H->bm_words);
		wi = (uint32_t)(key % H->bm_words);
		if (st_home(ld(state(H, sb))) != x)
			nrem = L.hn[slot];
	}
	/* the barrier above waited for every thread's stores; a superblock
	 * homed on another XCD also needs this XCD's L2 written back: one
	 * write-back and one counter update per wavefront (the count is summed
	 * bit-plane by bit-plane with ballots, which also works for a partial
	 * last wavefront) */
	if (__ballot(nrem != 0)) {
		uint32_t tot = 0;
		for (uint32_t b = 0; (1u << b) <= (uint32_t)BLOCK; b++)
			tot += (uint32_t)__popcll(__ballot((nrem >> b) & 1u)) << b;
		if (lane() == (uint32_t)__ffsll((unsigned long long)__ballot(1)) - 1u)
			fadd(stats(H) + KMA_STAT_REMOTE_FREE, (uint64_t)tot);
		release_l2();
	}
	if (owner) {
		uint32_t mask = L.hmask[slot];
		uint32_t old = f_and(bitmap(H, sb) + wi, ~mask);
		uint32_t n = __popc(old & mask);
		uint32_t dbl = L.hn[slot] - n;
		if (dbl)
			fadd(stats(H) + KMA_STAT_DOUBLE_FREE, (uint64_t)dbl);
		if (n)
			after_sub(H, sb, fsub(state(H, sb), (uint64_t)n) - n, n, x, ctx);
	}
	__syncthreads();
}

#endif /* KMA_WG_DEVICE_H */
