/*
 * kma_device.h - device-side kma_malloc / kma_free on the native gfx950 heap
 * layout (HIP).  Header-only so that user kernels can call it.
 *
 * Same call surface as the reference's device API (clheap.h:37-39): any lane
 * may call kma_malloc(heap, size) / kma_free(heap, ptr) on its own, with no
 * uniformity requirement; NULL on failure; kma_free(heap, NULL) is a no-op.
 *
 * Heap layout (one device buffer, all offsets from the header):
 *   kma_dheap header | stats | u64 state[nsb] | u32 pool_next[nsb] |
 *   u32 bitmap[nsb][bm_words] | u64 heads[nctx][KMA_HEADS_PER_CTX] |
 *   pool shards (128 B each) | u64 pbitmap[npages/64] | u8 plen[npages] |
 *   superblock data (nsb << sb_log2, SB-aligned) | page data.
 * Metadata lives outside the superblocks, so blocks are naturally aligned to
 * their (power-of-two) size and user data is never written by the allocator.
 *
 * Superblock state word (64-bit, one per superblock):
 *   bits  0..31  count: live blocks + in-flight reservations
 *   bits 32..36  class + 1 (0: superblock is free / in the pool)
 *   bit  37      owned: attached to a context head
 *   bits 38..40  home XCD: the XCD whose context attached it
 *   bit  41      listed: detached, partly free, in the partial index of (home, class)
 *   bits 42..63  generation, bumped each time the superblock returns to the pool
 *                or is attached again from the partial index
 * A wavefront group of n same-class lanes reserves with ONE atomicAdd(n) on
 * this word; the returned old value says (atomically) whether the superblock
 * is still the incarnation the group read from the context head (generation),
 * how many slots were left (count) and who must return it to the pool.
 * Whoever performs the atomic that leaves {count 0, not owned, class != 0}
 * returns the superblock to the pool with a CAS that also bumps the
 * generation.  Over-reservations are undone with atomicSub; a transiently
 * inflated count only makes a superblock look fuller than it is.
 *
 * Memory ordering.  Every allocator atomic is a relaxed agent-scope atomic:
 * RMWs are performed at the device-coherent level, and each protocol step
 * that must follow another consumes the earlier step's returned value.  An
 * agent-scope acquire/release on this part writes back / invalidates the
 * whole L2 of the XCD, so it is used only where a block's cache lines can
 * move to another XCD:
 *   - contexts and superblocks are homed on one XCD (contexts are chosen by
 *     the XCC id of the calling wavefront), so a block is normally reused on
 *     the XCD whose L2 holds its lines;
 *   - a free from another XCD first writes back its own L2 (release);
 *   - an emptied superblock goes to one of its home XCD's "dirty" pool
 *     shards, which only the home XCD pops from, with no write-back; other
 *     XCDs take from "clean" shards, which a home XCD refills in batches of
 *     up to 64 superblocks behind ONE L2 write-back, when asked (migration);
 *   - page frees always write back (pages are not homed).
 * A local free only waits for its own stores (s_waitcnt) before it clears
 * the block's bit.
 *
 * Context heads (64-bit): 0 empty, 1 refill in progress, otherwise
 * (gen << 32) | (sb + 2).  Pool shards are Treiber stacks with a 64-bit
 * {tag, top} head (ABA tag bumped on every pop and push); XCD x owns shards
 * [x*S, x*S + S) with S = npool/8: the first S/2 are its dirty shards, the
 * rest are clean.  Large requests (> largest class) take 1..64 contiguous pages from a
 * page bitmap (64-bit words, runs aligned to the next power of two).
 */
#ifndef KMA_DEVICE_H
#define KMA_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define KMA_NATIVE_MAGIC 0x4e414d4b31414d4bull /* "KMA1KMAN" */
#define KMA_HEADS_PER_CTX 32u
#define KMA_POOL_STRIDE 16u /* u64 words per pool shard (128 B) */
#define KMA_NXCD 8u         /* XCDs per MI355X */

typedef struct kma_dheap {
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
	uint64_t class_magic[32]; /* ceil(2^40 / size): exact block index */
	uint8_t class_for_log2[64];
	uint64_t off_psum;     /* page-window summaries: KMA_PAGE_LEVELS levels, then their top level */
	uint64_t off_spare;    /* context spare lines, 128 B per context */
	uint32_t spare_slots;  /* slots used per spare line (0..16) */
	uint32_t pad1;
	uint64_t off_pidx;     /* partial index: [8 XCDs][nclasses] blocks of pidx_words u64 */
	uint32_t pidx_words;   /* u64 per (XCD, class) block: levels 0, 1, 2 (see pidx) */
	uint32_t no_partial;   /* 1: no listing, no taking listed superblocks, no pressure routing */
} kma_dheap;

/* Per-heap counters (u64 at off_stats). */
#define KMA_STAT_DOUBLE_FREE 0
#define KMA_STAT_BAD_FREE 1
#define KMA_STAT_POOL_EMPTY 2
#define KMA_STAT_WAIT_RESERVE 3 /* a reservation kept finding its head busy (or retrying) for the watchdog period */
#define KMA_STAT_WAIT_CLAIM 4   /* a per-lane bitmap claim hit the watchdog */
#define KMA_STAT_WAIT_WGCLAIM 5 /* a group claim hit the watchdog, or a whole-superblock claim lost bits */
#define KMA_STAT_WAIT_POOL 6    /* a pool push/pop hit the watchdog */
#define KMA_STAT_REMOTE_FREE 7  /* frees issued from a non-home XCD */
#define KMA_STAT_MIGRATED 8     /* superblocks moved from dirty to clean shards in a kernel */
#define KMA_STAT_LISTED 9       /* detached superblocks put into a partial index */
#define KMA_STAT_REATTACHED 10  /* listed superblocks reused: attached to a head again, or grabbed */
#define KMA_NSTATS 11
/* The stats block is 256 B of counters and one 128-B line per XCD x holding
 * x's shard mask (bit k set when shard x*S + k may hold a superblock, see
 * pool_push / pool_pop_shard) and, in the next word, x's clean request
 * (non-zero: another XCD asks x to migrate dirty superblocks). */
#define KMA_STATS_BYTES (256u + 128u * 8u)
/* Page runs of k pages use an aligned window of 2^l >= k pages inside one
 * 64-page bitmap word; levels l = 0..6. */
#define KMA_PAGE_LEVELS 7u

/* Every wait in the native allocator gives up after this many ticks of the
 * constant 100 MHz wall clock (0.25 s).  Reaching it means the heap is
 * corrupt or a caller broke the protocol; the call then fails and bumps its
 * site's counter instead of spinning on. */
#define KMA_WAIT_TICKS 25000000ull

namespace kma_nat {

static constexpr uint64_t HEAD_EMPTY = 0;
static constexpr uint64_t HEAD_BUSY = 1;
static constexpr uint64_t ST_COUNT = 0xffffffffull;
static constexpr uint64_t ST_CLASS_SHIFT = 32;
static constexpr uint64_t ST_CLASS = 0x1full << ST_CLASS_SHIFT;
static constexpr uint64_t ST_OWNED = 1ull << 37;
static constexpr uint64_t ST_HOME_SHIFT = 38;
static constexpr uint64_t ST_LISTED = 1ull << 41;
static constexpr uint64_t ST_GEN_SHIFT = 42;
static constexpr uint32_t ST_GEN_MASK = 0x3fffffu;

/* Watchdog for retry loops: reads the wall clock every 256 retries only. */
struct waiter {
	uint64_t t0 = 0;
	uint32_t n = 0;
	__device__ __forceinline__ bool expired()
	{
		if ((++n & 255u) != 0)
			return false;
		uint64_t t = wall_clock64();
		if (t0 == 0) {
			t0 = t;
			return false;
		}
		return t - t0 > KMA_WAIT_TICKS;
	}
};

__device__ __forceinline__ uint32_t st_count(uint64_t s) { return (uint32_t)(s & ST_COUNT); }
__device__ __forceinline__ uint32_t st_cls1(uint64_t s) { return (uint32_t)((s & ST_CLASS) >> ST_CLASS_SHIFT); }
__device__ __forceinline__ uint32_t st_home(uint64_t s) { return (uint32_t)(s >> ST_HOME_SHIFT) & 7u; }
__device__ __forceinline__ uint32_t st_gen(uint64_t s) { return (uint32_t)(s >> ST_GEN_SHIFT); }

/* ---- relaxed agent-scope atomics ---- */
template <typename T>
__device__ __forceinline__ T ld(const T *p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st(T *p, T v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T cas(T *p, T cmp, T val)
{
	__hip_atomic_compare_exchange_strong(p, &cmp, val, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
					     __HIP_MEMORY_SCOPE_AGENT);
	return cmp;
}
template <typename T>
__device__ __forceinline__ T xchg(T *p, T v)
{
	return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T fadd(T *p, T v)
{
	return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T fsub(T *p, T v)
{
	return __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T f_or(T *p, T v)
{
	return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T f_and(T *p, T v)
{
	return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Wait until this wavefront's earlier memory operations have completed
 * (s_waitcnt vmcnt(0) expcnt(0) lgkmcnt(0)), with compiler barriers so no
 * access moves across it.  No cache maintenance: a completed store is in
 * this XCD's L2 and a completed atomic is performed at the coherent level. */
__device__ __forceinline__ void order_prior()
{
	__atomic_signal_fence(__ATOMIC_SEQ_CST);
	__builtin_amdgcn_s_waitcnt(0);
	__atomic_signal_fence(__ATOMIC_SEQ_CST);
}
/* Agent-scope release: also writes back this XCD's L2. */
__device__ __forceinline__ void release_l2()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

/* a - b if a > b, else 0.  Written without a compare-and-select: with a
 * select, `a >= b ? 0 : min(n, a - b)` was folded into a plain wrapping
 * subtraction inside large kernels by this toolchain's optimizer. */
__device__ __forceinline__ uint32_t sat_sub(uint32_t a, uint32_t b)
{
	return a - min(a, b);
}

/* XCD (XCC) the calling wavefront runs on: hwreg HW_REG_XCC_ID bits 0..3. */
__device__ __forceinline__ uint32_t xcc()
{
	return (uint32_t)__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & (KMA_NXCD - 1u);
}

/* ---- wavefront helpers ---- */
__device__ __forceinline__ uint32_t lane() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane()) - 1ull; }
__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t l)
{
	return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

/* ---- layout accessors ---- */
__device__ __forceinline__ char *base(const kma_dheap *H) { return (char *)H; }
__device__ __forceinline__ uint64_t *state(const kma_dheap *H, uint64_t sb) { return (uint64_t *)(base(H) + H->off_state) + sb; }
__device__ __forceinline__ uint32_t *pnext(const kma_dheap *H, uint64_t sb) { return (uint32_t *)(base(H) + H->off_next) + sb; }
__device__ __forceinline__ uint32_t *bitmap(const kma_dheap *H, uint64_t sb) { return (uint32_t *)(base(H) + H->off_bitmap) + sb * H->bm_words; }
__device__ __forceinline__ uint64_t *head(const kma_dheap *H, uint32_t ctx, uint32_t c) { return (uint64_t *)(base(H) + H->off_heads) + (uint64_t)ctx * KMA_HEADS_PER_CTX + c; }
__device__ __forceinline__ uint64_t *pool(const kma_dheap *H, uint32_t s) { return (uint64_t *)(base(H) + H->off_pool) + (uint64_t)s * KMA_POOL_STRIDE; }
__device__ __forceinline__ uint64_t *pbits(const kma_dheap *H) { return (uint64_t *)(base(H) + H->off_pbitmap); }
__device__ __forceinline__ uint8_t *plen(const kma_dheap *H) { return (uint8_t *)(base(H) + H->off_plen); }
__device__ __forceinline__ char *sbdata(const kma_dheap *H, uint64_t sb) { return base(H) + H->off_data + (sb << H->sb_log2); }
__device__ __forceinline__ uint64_t *stats(const kma_dheap *H) { return (uint64_t *)(base(H) + H->off_stats); }
__device__ __forceinline__ uint64_t *pmask(const kma_dheap *H, uint32_t x) { return (uint64_t *)(base(H) + H->off_stats + 256u + 128u * x); }
__device__ __forceinline__ uint64_t *creq(const kma_dheap *H, uint32_t x) { return pmask(H, x) + 1; }
/* wall-clock tick until which XCD x's pops skip the search beyond its own
 * spares and shards (set by a search that found nothing) */
__device__ __forceinline__ uint64_t *exhausted(const kma_dheap *H, uint32_t x) { return pmask(H, x) + 2; }
/* non-zero from a pool search of XCD x that found nothing to the next pop
 * of x that finds a superblock: x's free capacity is in partly free ones */
__device__ __forceinline__ uint64_t *pressure(const kma_dheap *H, uint32_t x) { return pmask(H, x) + 3; }
__device__ __forceinline__ uint64_t psum_words(const kma_dheap *H) { return (((H->npages + 63) >> 6) + 63) >> 6; }
__device__ __forceinline__ uint64_t *psum(const kma_dheap *H, uint32_t l) { return (uint64_t *)(base(H) + H->off_psum) + l * psum_words(H); }
__device__ __forceinline__ uint64_t psum2_words(const kma_dheap *H) { return (psum_words(H) + 63) >> 6; }
__device__ __forceinline__ uint64_t *psum2(const kma_dheap *H, uint32_t l) { return psum(H, KMA_PAGE_LEVELS) + l * psum2_words(H); }

__device__ __forceinline__ void bump_stat(const kma_dheap *H, int k)
{
	fadd(stats(H) + k, (uint64_t)1);
}

/* Size -> class: the first class holding 2^(ceil_log2(size)-1)+1, then walk
 * up (at most the number of classes within one octave). */
__device__ __forceinline__ int class_of(const kma_dheap *H, uint32_t size)
{
	uint32_t e = size <= 1 ? 0 : 32 - __clz(size - 1);
	int c = H->class_for_log2[e];
	while (H->class_size[c] < size)
		c++;
	return c;
}

/* Contexts of XCD x are [x*P, x*P + P), P = nctx/8.  Within its XCD, a
 * wavefront maps to a context by (work-group, wave) so that a few nearby
 * wavefronts share each context's per-class superblocks. */
__device__ __forceinline__ uint32_t ctx_of(const kma_dheap *H, uint32_t x)
{
	uint32_t per = H->nctx / KMA_NXCD;
	uint32_t wpb = (blockDim.x * blockDim.y * blockDim.z + 63u) >> 6;
	uint32_t tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
	uint32_t bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
	return x * per + ((bid / KMA_NXCD) * wpb + (tid >> 6)) % per;
}

/* Context of a whole work-group (work-group collective API). */
__device__ __forceinline__ uint32_t wg_ctx_of(const kma_dheap *H, uint32_t x)
{
	uint32_t per = H->nctx / KMA_NXCD;
	uint32_t bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
	return x * per + (bid / KMA_NXCD) % per;
}

/* ---- superblock pool: sharded Treiber stacks, 64-bit {tag:32, top:32} ----
 * XCD x owns shards [x*S, x*S + S), S = npool/8 <= 64: the first D = S/2
 * are its dirty shards (only x pops them), the other S - D are clean.
 * A superblock emptied on its home XCD x goes to a dirty shard of x with no
 * cache maintenance: x's L2 may still hold dirty lines of its blocks, which
 * must not be written back over another XCD's later use.  Another XCD gets
 * such superblocks only through a migration run on x (migrate_dirty): pop a
 * batch, write back x's L2 once, push the batch onto a clean shard.  An XCD
 * that finds no clean superblock anywhere sets the clean request of each XCD
 * holding dirty ones; those XCDs migrate when they next pop a superblock,
 * and kma_trim (after a kernel boundary) empties every dirty shard.
 * pmask(x) bit k is a hint that shard x*S + k is non-empty, updated only
 * when a shard goes empty <-> non-empty:
 *   - a push that found the shard empty sets the bit after its CAS;
 *   - a pop that left the shard empty (or found it empty) clears the bit,
 *     waits for that to complete and re-reads the head, setting the bit
 *     again if a push got in.
 * So a non-empty shard's bit is clear only while a push onto it is in
 * flight; a set bit on an empty shard costs one wasted probe. */
__device__ __forceinline__ uint32_t shards_per_xcd(const kma_dheap *H) { return H->npool / KMA_NXCD; }
__device__ __forceinline__ uint32_t dirty_per_xcd(const kma_dheap *H) { return shards_per_xcd(H) / 2u; }
__device__ __forceinline__ uint32_t dirty_shard(const kma_dheap *H, uint32_t x, uint32_t k)
{
	return x * shards_per_xcd(H) + k % dirty_per_xcd(H);
}
__device__ __forceinline__ uint32_t clean_shard(const kma_dheap *H, uint32_t x, uint32_t k)
{
	uint32_t S = shards_per_xcd(H), D = dirty_per_xcd(H);
	return x * S + D + k % (S - D);
}
__device__ __forceinline__ uint64_t shard_bit(const kma_dheap *H, uint32_t shard)
{
	return (uint64_t)1 << (shard % shards_per_xcd(H));
}

/* The shard went (or was found) empty: clear its hint unless a push got in.
 * A hint already clear is left alone (read first: under exhaustion many
 * poppers find the same shard empty, and an RMW each would queue them all
 * on the XCD's mask word). */
__device__ inline void shard_emptied(const kma_dheap *H, uint32_t shard)
{
	uint64_t *mp = pmask(H, shard / shards_per_xcd(H));
	uint64_t bit = shard_bit(H, shard);
	if (!(ld(mp) & bit))
		return;
	f_and(mp, ~bit);
	order_prior();
	if ((uint32_t)ld(pool(H, shard)) != 0)
		f_or(mp, bit);
}

/* Push the chain first -> ... -> last (n superblocks linked through pnext)
 * onto a shard.  A push never gives up: a failed CAS means another push or
 * pop succeeded, so the pool as a whole progresses; back-off spreads the
 * retries. */
__device__ inline void pool_push_chain(const kma_dheap *H, uint32_t shard, uint32_t first, uint32_t last,
				       uint32_t n)
{
	uint64_t *hp = pool(H, shard);
	uint64_t h = ld(hp);
	uint32_t bo = 1;
	for (;;) {
		/* the link must be performed before the head CAS publishes it */
		xchg(pnext(H, last), (uint32_t)h);
		order_prior();
		uint64_t nh = (((h >> 32) + 1ull) << 32) | (uint64_t)(first + 1u);
		uint64_t seen = cas(hp, h, nh);
		if (seen == h)
			break;
		h = seen;
		for (uint32_t i = 0; i < bo; i++)
			__builtin_amdgcn_s_sleep(1);
		bo = bo < 2u ? 2u * bo : 2u;
	}
	if ((uint32_t)h == 0)
		f_or(pmask(H, shard / shards_per_xcd(H)), shard_bit(H, shard));
	fadd(hp + 1, (uint64_t)n);
}

__device__ __forceinline__ void pool_push(const kma_dheap *H, uint32_t shard, uint32_t sb)
{
	pool_push_chain(H, shard, sb, sb, 1u);
}

__device__ inline bool pool_pop_shard(const kma_dheap *H, uint32_t shard, uint32_t *out)
{
	uint64_t *hp = pool(H, shard);
	uint64_t h = ld(hp);
	waiter wt;
	uint32_t bo = 1;
	for (;;) {
		if (wt.expired()) {
			bump_stat(H, KMA_STAT_WAIT_POOL);
			return false;
		}
		uint32_t top = (uint32_t)h;
		if (top == 0) {
			shard_emptied(H, shard);
			return false;
		}
		uint32_t nx = ld(pnext(H, top - 1u));
		uint64_t nh = (((h >> 32) + 1ull) << 32) | (uint64_t)nx;
		uint64_t seen = cas(hp, h, nh);
		if (seen == h) {
			fsub(hp + 1, (uint64_t)1);
			if (nx == 0)
				shard_emptied(H, shard);
			*out = top - 1u;
			return true;
		}
		h = seen;
		for (uint32_t i = 0; i < bo; i++)
			__builtin_amdgcn_s_sleep(1);
		bo = bo < 2u ? 2u * bo : 2u;
	}
}

/* Pop from the shards of XCD y whose hint bits are in `bits`, starting at
 * the bit after position `start`. */
__device__ inline bool pool_pop_bits(const kma_dheap *H, uint32_t y, uint64_t bits, uint32_t start, uint32_t *out)
{
	uint32_t S = shards_per_xcd(H);
	uint32_t r = start % S;
	uint64_t hi = bits & (~0ull << r), lo = bits & ~(~0ull << r);
	for (int pass = 0; pass < 2; pass++) {
		uint64_t m = pass ? lo : hi;
		while (m) {
			uint32_t k = (uint32_t)__ffsll((unsigned long long)m) - 1u;
			m &= m - 1ull;
			if (pool_pop_shard(H, y * S + k, out))
				return true;
		}
	}
	return false;
}

/* ---- context spares ----
 * Each context has a 128-B line of 16 spare slots (u32 sb + 1, 0: empty;
 * the first spare_slots of them are used): empty superblocks the context keeps
 * for its next refill, so that most superblock turnover touches only the
 * context's own line instead of the shared pool shards.  A spare is in the
 * same state as a pooled superblock (count 0, no class, generation bumped)
 * and, like a dirty shard's, is reused only on its context's XCD: a pool
 * pop that finds its own XCD's shards empty also looks into a few other
 * contexts' lines of that XCD, a migration drains the XCD's lines, and
 * kma_trim empties them all.  Capacity is at most a quarter of the heap. */
__device__ __forceinline__ uint32_t *spare(const kma_dheap *H, uint32_t ctx)
{
	return (uint32_t *)(base(H) + H->off_spare) + (uint64_t)ctx * 32u;
}


/* Keep empty superblock sb as a spare of context ctx: a CAS into slot
 * sb mod n, then into the slot half a line away; false if both are taken. */
__device__ __forceinline__ bool spare_put(const kma_dheap *H, uint32_t ctx, uint32_t sb)
{
	uint32_t n = H->spare_slots;
	if (n == 0u)
		return false;
	uint32_t *sp = spare(H, ctx);
	uint32_t k = sb % n;
	if (cas(sp + k, 0u, sb + 1u) == 0u)
		return true;
	return n > 1u && cas(sp + (k + n / 2u) % n, 0u, sb + 1u) == 0u;
}

/* Occupied slots of ctx's line (bit k: slot k), read in one round trip. */
__device__ __forceinline__ uint32_t spare_occupied(const kma_dheap *H, uint32_t ctx)
{
	uint32_t n = H->spare_slots, occ = 0;
	const uint64_t *sp = (const uint64_t *)spare(H, ctx);
#pragma unroll
	for (uint32_t i = 0; i < 8u; i++) {
		if (2u * i < n) {
			uint64_t v = ld(sp + i);
			occ |= ((uint32_t)v != 0u ? 1u : 0u) << (2u * i);
			occ |= ((uint32_t)(v >> 32) != 0u ? 1u : 0u) << (2u * i + 1u);
		}
	}
	return occ & ((1u << n) - 1u);
}

/* Swap out an occupied slot of ctx's line (occ: as read), preferring slots
 * from `start` on; false if every candidate was emptied meanwhile. */
__device__ inline bool spare_take(const kma_dheap *H, uint32_t ctx, uint32_t occ, uint32_t start, uint32_t *out)
{
	uint32_t n = H->spare_slots;
	uint32_t *sp = spare(H, ctx);
	uint32_t r = n ? start % n : 0u;
	while (occ) {
		uint32_t rot = r ? ((occ >> r) | (occ << (n - r))) & ((1u << n) - 1u) : occ;
		uint32_t k = (__ffs(rot) - 1u + r) % n;
		occ &= ~(1u << k);
		uint32_t v = xchg(sp + k, 0u);
		if (v) {
			*out = v - 1u;
			return true;
		}
	}
	return false;
}

/* Make up to KMA_MIGRATE_MAX of XCD x's dirty superblocks usable by every
 * XCD (runs on x, one lane): pop them, chaining them through pnext, write
 * back x's L2 once, push the chain onto one of x's clean shards.  Every
 * store to a popped superblock completed in x's L2 (or was written back by a
 * remote freer) before its push, which the pop observed, so the one
 * write-back covers the whole batch. */
#define KMA_MIGRATE_MAX 64u
#define KMA_MIGRATE_CTX 32u
__device__ inline void migrate_dirty(const kma_dheap *H, uint32_t x, uint32_t ctx)
{
	uint32_t S = shards_per_xcd(H), D = dirty_per_xcd(H);
	uint64_t mask = ld(pmask(H, x)) & ((1ull << D) - 1ull);
	uint32_t first = 0, last = 0, n = 0, sb;
	for (uint32_t i = 0; i < D && n < KMA_MIGRATE_MAX; i++) {
		uint32_t k = (ctx + i) % D;
		if (!((mask >> k) & 1ull))
			continue;
		while (n < KMA_MIGRATE_MAX && pool_pop_shard(H, x * S + k, &sb)) {
			xchg(pnext(H, sb), first);
			if (n == 0)
				last = sb;
			first = sb + 1u;
			n++;
		}
	}
	/* then the spares of up to KMA_MIGRATE_CTX of x's contexts (a rare
	 * path: one line at a time keeps the callers' register budget) */
	uint32_t per = H->nctx / KMA_NXCD;
	for (uint32_t i = 0; i < per && i < KMA_MIGRATE_CTX && n < KMA_MIGRATE_MAX && H->spare_slots; i++) {
		uint32_t c2 = x * per + (ctx + i) % per;
		uint32_t *sp = spare(H, c2);
		for (uint32_t m = spare_occupied(H, c2); m && n < KMA_MIGRATE_MAX; m &= m - 1u) {
			uint32_t v = xchg(sp + (__ffs(m) - 1u), 0u);
			if (!v)
				continue;
			sb = v - 1u;
			xchg(pnext(H, sb), first);
			if (n == 0)
				last = sb;
			first = sb + 1u;
			n++;
		}
	}
	if (n == 0)
		return;
	release_l2();
	pool_push_chain(H, clean_shard(H, x, ctx), first - 1u, last, n);
	fadd(stats(H) + KMA_STAT_MIGRATED, (uint64_t)n);
}

/* Run a migration if another XCD asked XCD x for one (rq: x's request word
 * as read by the caller); the lane that clears the request does it. */
__device__ __forceinline__ void serve_creq(const kma_dheap *H, uint32_t x, uint32_t ctx, uint64_t rq)
{
	if (rq && xchg(creq(H, x), (uint64_t)0))
		migrate_dirty(H, x, ctx);
}

/* Pop for XCD x and context ctx: ctx's spares, x's clean shards, x's dirty
 * shards, a few other contexts' spares on x, the other XCDs' clean shards
 * (clean before dirty: popping the dirty shards that this XCD's releases
 * are pushing onto measured slower).  `salt` spreads concurrent callers of
 * one context over slots and shards.  A pop reads only its own
 * XCD's mask/request line until its own shards are exhausted (every pop of
 * the device reading all eight masks made them hot), then the other masks
 * in one round trip.  With no clean superblock anywhere, a caller that
 * newly requests a migration from the XCDs holding dirty ones waits up to
 * KMA_CREQ_POLLS short polls for one (serving requests aimed at its own XCD
 * meanwhile); requests already pending mean those XCDs have not answered,
 * and the caller fails at once.  (Never waiting was 11-18 % faster on the
 * oversubscribed request stream, but returned NULLs from heaps with room on
 * other XCDs.)
 * KMA_STEAL_CTX other contexts' spare lines are looked into per search (8
 * cost 4-8 % over 2 on that stream, where most searches fail). */
#ifndef KMA_CREQ_POLLS
#define KMA_CREQ_POLLS 1u
#endif
#ifndef KMA_STEAL_CTX
#define KMA_STEAL_CTX 2u
#endif
/* After a search that found no superblock anywhere, XCD x's pops give up
 * after the caller's spares, x's own shards and a few other contexts'
 * spares on x, for this many wall-clock ticks (10 ns each): under
 * exhaustion every malloc would otherwise repeat the whole search (every
 * XCD's masks and clean shards, a migration request), and those searches,
 * all at once, queue on the same few words.  (200 us: 6-11 % faster than
 * 50 us on the oversubscribed request stream, with the same NULL count.) */
#ifndef KMA_EXHAUSTED_TICKS
#define KMA_EXHAUSTED_TICKS 20000u
#endif
__device__ inline bool pool_search(const kma_dheap *H, uint32_t x, uint32_t ctx, uint32_t salt, uint32_t *out)
{
	uint32_t S = shards_per_xcd(H), D = dirty_per_xcd(H);
	uint64_t all = S == 64 ? ~0ull : ((1ull << S) - 1ull);
	uint64_t dirty = (1ull << D) - 1ull, clean = all & ~dirty;
	uint32_t sh = ctx + salt;
	uint64_t mx = ld(pmask(H, x));
	uint64_t rq = ld(creq(H, x));
	uint64_t ex = ld(exhausted(H, x));
	uint32_t occ = H->spare_slots ? spare_occupied(H, ctx) : 0u;
	serve_creq(H, x, ctx, rq);
	if (occ && spare_take(H, ctx, occ, salt, out))
		return true;
	if (pool_pop_bits(H, x, mx & clean, D + sh % (S - D), out))
		return true;
	if (pool_pop_bits(H, x, mx & dirty, sh % D, out))
		return true;
	/* other contexts' spares on this XCD, a few per call */
	uint32_t per = H->nctx / KMA_NXCD;
	if (H->spare_slots && per > 1u) {
		uint32_t o[KMA_STEAL_CTX];
		uint32_t c0 = (uint32_t)wall_clock64() + salt;
#pragma unroll
		for (uint32_t i = 0; i < KMA_STEAL_CTX; i++)
			o[i] = i < per - 1u ? spare_occupied(H, x * per + (ctx - x * per + 1u + (c0 + i) % (per - 1u)) % per) : 0u;
#pragma unroll
		for (uint32_t i = 0; i < KMA_STEAL_CTX; i++)
			if (o[i] && spare_take(H, x * per + (ctx - x * per + 1u + (c0 + i) % (per - 1u)) % per, o[i], salt, out))
				return true;
	}
	if (ex && (uint64_t)wall_clock64() < ex)
		return false;
	/* m[j]: the mask of XCD (x + j) mod 8 (constant indices keep the
	 * array in registers) */
	uint64_t m[KMA_NXCD];
	for (uint32_t poll = 0;; poll++) {
#pragma unroll
		for (uint32_t j = 0; j < KMA_NXCD; j++)
			m[j] = ld(pmask(H, (x + j) & (KMA_NXCD - 1u)));
		if (poll > 0) {
			serve_creq(H, x, ctx, ld(creq(H, x)));
			if (pool_pop_bits(H, x, m[0] & clean, D + sh % (S - D), out))
				return true;
			if (pool_pop_bits(H, x, m[0] & dirty, sh % D, out))
				return true;
		}
#pragma unroll
		for (uint32_t j = 1; j < KMA_NXCD; j++)
			if (pool_pop_bits(H, (x + j) & (KMA_NXCD - 1u), m[j] & clean, D + sh % (S - D), out))
				return true;
		if (poll == 0) {
			/* ask the XCDs holding dirty superblocks for a migration; a
			 * request already pending (read first, so that failing
			 * callers do not all RMW the same words) means they have not
			 * answered yet, and this caller fails at once */
			uint32_t asked = 0;
#pragma unroll
			for (uint32_t j = 1; j < KMA_NXCD; j++) {
				uint64_t *rw = creq(H, (x + j) & (KMA_NXCD - 1u));
				if ((m[j] & dirty) && ld(rw) == 0 && !xchg(rw, (uint64_t)1))
					asked++;
			}
			if (asked == 0)
				break;
		}
		if (poll == KMA_CREQ_POLLS)
			break;
		__builtin_amdgcn_s_sleep(32);
	}
	st(exhausted(H, x), (uint64_t)(wall_clock64() + KMA_EXHAUSTED_TICKS));
	return false;
}

/* pool_search, keeping XCD x's pressure word: set by a search that found
 * nothing, cleared by the next one that found a superblock (read with the
 * search's own first loads, so a pop under no pressure adds no round trip) */
__device__ inline bool pool_pop(const kma_dheap *H, uint32_t x, uint32_t ctx, uint32_t salt, uint32_t *out)
{
	const uint64_t pr = ld(pressure(H, x));
	const bool ok = pool_search(H, x, ctx, salt, out);
	if (ok && pr)
		st(pressure(H, x), (uint64_t)0);
	else if (!ok && !pr)
		st(pressure(H, x), (uint64_t)1);
	return ok;
}

/* ---- partial index ----
 * A superblock detached from its head (full when it was detached) whose
 * frees leave at least an eighth of it free is "listed": the LISTED bit of
 * its state word is set and its bit is set in the partial index of (home
 * XCD, class).  A head refill whose pool search finds no empty superblock
 * takes one from its own XCD's index of its class and attaches it again
 * (generation bumped), so capacity freed block by block is reused before
 * the superblock drains (a detached superblock is otherwise reachable only
 * once it is empty).  The index of one (XCD, class) is three levels of u64
 * words: level 0 has one bit per superblock, level 1 one bit per level-0
 * word ("may be non-zero"), level 2 one bit per level-1 word.  A put sets
 * the superblock's bit and, if that word was zero, the word's bit one level
 * up (and so on); whoever leaves a word zero (or finds it zero) clears its
 * bit one level up, re-reads the word and sets the bit again if a put came
 * in -- the protocol of the pool's shard masks and the page summaries, so a
 * set bit is a hint and a non-zero word always has its bit set once puts
 * have finished.  A take walks down set bits (from a salted position) and
 * claims a superblock with one atomicAnd on its level-0 word; the free that
 * empties a listed superblock clears its bit the same way, and whichever of
 * the two clears it owns the superblock.  Invariants: LISTED is set only
 * while the superblock is not owned and has a class, and then its bit is
 * set or being set (by its lister); whoever clears the bit clears LISTED.
 * Indexes are per XCD for the same reason as dirty shards: a block is
 * reused on the XCD whose L2 may hold its lines. */
__device__ __forceinline__ uint32_t pidx_w0(const kma_dheap *H) { return (uint32_t)((H->nsb + 63u) >> 6); }
__device__ __forceinline__ uint32_t pidx_w1(const kma_dheap *H) { return (pidx_w0(H) + 63u) >> 6; }
__device__ __forceinline__ uint64_t *pidx(const kma_dheap *H, uint32_t x, uint32_t c)
{
	return (uint64_t *)(base(H) + H->off_pidx) + ((uint64_t)x * H->nclasses + c) * H->pidx_words;
}
/* First word of level lv (0, 1, 2) in an index block. */
__device__ __forceinline__ uint32_t pidx_level(const kma_dheap *H, uint32_t lv)
{
	return lv == 0u ? 0u : lv == 1u ? pidx_w0(H) : pidx_w0(H) + pidx_w1(H);
}
/* Free slots a detached superblock needs before it is listed: an eighth of
 * it, or a sixteenth while its XCD is under pressure (the free capacity of
 * a nearly full heap is a few blocks per superblock; listing superblocks
 * with fewer would make every take of the index yield a block or two). */
__device__ __forceinline__ uint32_t part_room(uint32_t cap) { return cap >= 8u ? cap / 8u : 1u; }
__device__ __forceinline__ uint32_t part_room_pressed(uint32_t cap) { return cap >= 16u ? cap / 16u : 1u; }

/* Set bit i of level lv; a word that was zero sets its own bit one level up
 * (a summary bit that reads set already is not written again: the few
 * summary words are the index's hottest). */
__device__ inline void pidx_set(const kma_dheap *H, uint64_t *P, uint32_t lv, uint32_t i)
{
	for (; lv < 3u; lv++, i >>= 6) {
		uint64_t *w = P + pidx_level(H, lv) + (i >> 6);
		const uint64_t b = (uint64_t)1 << (i & 63u);
		if (lv > 0u && (ld(w) & b))
			return;
		if (f_or(w, b) != 0)
			return;
	}
}

/* Word i of level lv (0 or 1) was seen zero (after the caller's RMW, or by a
 * load): clear its bit one level up, re-read it and set the bit again if a
 * put came in; a level-1 word left zero by that clear continues one level
 * up. */
__device__ inline void pidx_emptied(const kma_dheap *H, uint64_t *P, uint32_t lv, uint32_t i)
{
	for (; lv < 2u; lv++, i >>= 6) {
		uint64_t *up = P + pidx_level(H, lv + 1u) + (i >> 6);
		const uint64_t b = (uint64_t)1 << (i & 63u);
		const uint64_t o = f_and(up, ~b);
		if (ld(P + pidx_level(H, lv) + i) != 0) {
			pidx_set(H, P, lv + 1u, i);
			return;
		}
		if ((o & ~b) != 0)
			return;
	}
}

/* Lowest set bit of m after rotating it right by r. */
__device__ __forceinline__ uint32_t pick_bit(uint64_t m, uint32_t r)
{
	r &= 63u;
	const uint64_t rot = r ? (m >> r) | (m << (64u - r)) : m;
	return (uint32_t)(__ffsll((unsigned long long)rot) - 1 + (int)r) & 63u;
}

__device__ __forceinline__ void pidx_put(const kma_dheap *H, uint32_t x, uint32_t c, uint32_t sb)
{
	pidx_set(H, pidx(H, x, c), 0u, sb);
}

/* Take any listed superblock of (x, c) out of the index; false if none. */
__device__ inline bool pidx_take(const kma_dheap *H, uint32_t x, uint32_t c, uint32_t salt, uint32_t *out)
{
	if (H->no_partial)
		return false;
	uint64_t *P = pidx(H, x, c);
	const uint32_t w1 = pidx_w1(H), w2 = (w1 + 63u) >> 6;
	const uint32_t L1 = pidx_level(H, 1u), L2 = pidx_level(H, 2u);
	for (uint32_t a = 0; a < w2; a++) {
		const uint32_t i2 = (a + salt) % w2;
		for (uint64_t t2 = ld(P + L2 + i2); t2;) {
			const uint32_t j2 = pick_bit(t2, salt);
			t2 &= ~((uint64_t)1 << j2);
			const uint32_t i1 = i2 * 64u + j2;
			uint64_t t1 = ld(P + L1 + i1);
			if (!t1)
				pidx_emptied(H, P, 1u, i1);
			while (t1) {
				const uint32_t j1 = pick_bit(t1, salt >> 6);
				t1 &= ~((uint64_t)1 << j1);
				const uint32_t i0 = i1 * 64u + j1;
				uint64_t t0 = ld(P + i0);
				while (t0) {
					const uint64_t b = (uint64_t)1 << pick_bit(t0, salt >> 12);
					const uint64_t o = f_and(P + i0, ~b);
					if (o & b) {
						if ((o & ~b) == 0)
							pidx_emptied(H, P, 0u, i0);
						*out = i0 * 64u + (uint32_t)(__ffsll((unsigned long long)b) - 1);
						return true;
					}
					t0 = o & ~b;
				}
				pidx_emptied(H, P, 0u, i0);
			}
		}
	}
	return false;
}

/* Clear sb's own bit in (x, c)'s index; false if it was clear (a refill
 * took it, or its lister has not set it yet). */
__device__ inline bool pidx_remove(const kma_dheap *H, uint32_t x, uint32_t c, uint32_t sb)
{
	uint64_t *P = pidx(H, x, c);
	const uint64_t b = (uint64_t)1 << (sb & 63u);
	const uint64_t o = f_and(P + (sb >> 6), ~b);
	if (!(o & b))
		return false;
	if ((o & ~b) == 0)
		pidx_emptied(H, P, 0u, sb >> 6);
	return true;
}

/* XCD x's last pool search found no empty superblock (its refills are
 * taking listed superblocks). */
__device__ __forceinline__ bool under_pressure(const kma_dheap *H, uint32_t x)
{
	return !H->no_partial && ld(pressure(H, x)) != 0;
}

/* Called by whoever produced state `s` == {count 0, not owned, not listed,
 * class set}, running on XCD x.  A local release keeps the superblock in
 * the releasing context's spare line unless the XCD is under pressure:
 * then every context of the XCD must find it, in the XCD's dirty shard. */
__device__ inline void try_release(const kma_dheap *H, uint32_t sb, uint64_t s, uint32_t x, uint32_t ctx)
{
	uint32_t g = st_gen(s);
	for (;;) {
		uint64_t fresh = (uint64_t)((g + 1u) & ST_GEN_MASK) << ST_GEN_SHIFT;
		uint64_t seen = cas(state(H, sb), s, fresh);
		if (seen == s) {
			/* a local release leaves this XCD's L2 as it is; a remote
			 * freer has written its own L2 back (kma_free) */
			uint32_t home = st_home(s);
			if (home != x || under_pressure(H, x) || !spare_put(H, ctx, sb))
				pool_push(H, dirty_shard(H, home, ctx), sb);
			return;
		}
		s = seen;
		if (st_count(s) != 0 || (s & (ST_OWNED | ST_LISTED)) || st_cls1(s) == 0 || st_gen(s) != g)
			return;
	}
}

/* Set listed superblock sb's index bit (s: its state, LISTED set).  A free
 * that empties it before the bit is set finds no bit and leaves it to this
 * call: the state is read again after the put, and if this incarnation
 * (same generation, still listed, not owned) has emptied, its bit is taken
 * back here and it is released (a refill, or that free, may have taken the
 * bit first: then nothing is left to do). */
__device__ inline void put_listed(const kma_dheap *H, uint32_t sb, uint64_t s, uint32_t x, uint32_t ctx)
{
	const uint32_t c = st_cls1(s) - 1u;
	pidx_put(H, st_home(s), c, sb);
	const uint64_t now = ld(state(H, sb));
	if (st_count(now) == 0 && (now & ~ST_COUNT) == (s & ~ST_COUNT) && pidx_remove(H, st_home(s), c, sb)) {
		const uint64_t r = f_and(state(H, sb), ~ST_LISTED) & ~ST_LISTED;
		if (st_count(r) == 0 && !(r & ST_OWNED))
			try_release(H, sb, r, x, ctx);
	}
}

/* List detached superblock sb (state s as last seen: not owned, class set,
 * count in 1..lim): set LISTED with a CAS while that still holds, then set
 * its bit in the index of (home, class) (put_listed). */
__device__ inline void try_list(const kma_dheap *H, uint32_t sb, uint64_t s, uint32_t lim, uint32_t x, uint32_t ctx)
{
	if (H->no_partial)
		return;
	const uint32_t c1 = st_cls1(s);
	for (;;) {
		if ((s & (ST_OWNED | ST_LISTED)) || st_cls1(s) != c1 || st_count(s) == 0 || st_count(s) > lim)
			return;
		const uint64_t seen = cas(state(H, sb), s, s | ST_LISTED);
		if (seen == s)
			break;
		s = seen;
	}
	bump_stat(H, KMA_STAT_LISTED);
	put_listed(H, sb, s | ST_LISTED, x, ctx);
}

/* A listed superblock just taken out of the index by a partial grab
 * (LISTED still set) goes back in if it still has room to offer (an eighth
 * of it free, or a sixteenth while its XCD is under pressure: press, as the
 * caller read it); otherwise its LISTED bit is cleared, and its frees list
 * it again when they cross to that much room. */
__device__ inline void relist(const kma_dheap *H, uint32_t sb, uint32_t x, uint32_t ctx, bool press)
{
	uint64_t s = ld(state(H, sb));
	for (;;) {
		const uint32_t cap = H->class_cap[st_cls1(s) - 1u], cnt = st_count(s);
		if (cnt <= cap - part_room(cap) || (press && cnt <= cap - part_room_pressed(cap))) {
			put_listed(H, sb, s, x, ctx);
			return;
		}
		const uint64_t seen = cas(state(H, sb), s, s & ~ST_LISTED);
		if (seen == s)
			return;
		s = seen;
	}
}

/* The free that emptied listed superblock sb (state s: count 0, listed):
 * clear its index bit and release it; if the bit was clear, a refill took it
 * (and attaches it) or its lister has not set it yet (the lister then finds
 * it empty and releases it, see try_list). */
__device__ inline void unlist_release(const kma_dheap *H, uint32_t sb, uint64_t s, uint32_t x, uint32_t ctx)
{
	if (!pidx_remove(H, st_home(s), st_cls1(s) - 1u, sb))
		return;
	/* out of the index: only a reservation still in flight against the head
	 * it once was can change the count now, and then its free releases it */
	s = f_and(state(H, sb), ~ST_LISTED) & ~ST_LISTED;
	if (st_count(s) == 0 && !(s & ST_OWNED))
		try_release(H, sb, s, x, ctx);
}

/* After a count decrement of m on superblock sb left state ns: release it
 * if it emptied with no owner; list it, while detached, when it crosses to
 * an eighth free, or at any free that leaves a sixteenth free while its
 * home XCD's pool is exhausted. */
__device__ __forceinline__ void after_sub(const kma_dheap *H, uint32_t sb, uint64_t ns, uint32_t m, uint32_t x,
					 uint32_t ctx)
{
	if ((ns & ST_OWNED) || st_cls1(ns) == 0)
		return;
	const uint32_t cnt = st_count(ns);
	if (cnt == 0) {
		if (ns & ST_LISTED)
			unlist_release(H, sb, ns, x, ctx);
		else
			try_release(H, sb, ns, x, ctx);
		return;
	}
	if (ns & ST_LISTED)
		return;
	const uint32_t cap = H->class_cap[st_cls1(ns) - 1u], lim = cap - part_room(cap);
	if (cnt <= lim && cnt + m > lim)
		try_list(H, sb, ns, lim, x, ctx);
	else if (cnt > lim && cnt <= cap - part_room_pressed(cap) && under_pressure(H, st_home(ns)))
		try_list(H, sb, ns, cap - part_room_pressed(cap), x, ctx);
}

/* After the owned bit was cleared (state s): release or list. */
__device__ __forceinline__ void after_detach(const kma_dheap *H, uint32_t sb, uint64_t s, uint32_t x, uint32_t ctx)
{
	if (st_cls1(s) == 0)
		return;
	const uint32_t cap = H->class_cap[st_cls1(s) - 1u], cnt = st_count(s);
	if (cnt == 0)
		try_release(H, sb, s, x, ctx);
	else if (cnt <= cap - part_room(cap))
		try_list(H, sb, s, cap - part_room(cap), x, ctx);
	else if (cnt <= cap - part_room_pressed(cap) && under_pressure(H, st_home(s)))
		try_list(H, sb, s, cap - part_room_pressed(cap), x, ctx);
}

__device__ __forceinline__ void undo(const kma_dheap *H, uint32_t sb, uint32_t m, uint32_t x, uint32_t ctx)
{
	uint64_t s = fsub(state(H, sb), (uint64_t)m) - m;
	after_sub(H, sb, s, m, x, ctx);
}

/* Quiescent heap only (kma_trim): release every listed superblock of
 * level-0 word i0 of (x, c)'s index that holds no live block (the summary
 * levels are rebuilt afterwards by pidx_rebuild). */
__device__ inline void trim_pidx_word(const kma_dheap *H, uint32_t x, uint32_t c, uint32_t i0)
{
	uint64_t *w = pidx(H, x, c) + i0;
	for (uint64_t m = ld(w); m; m &= m - 1u) {
		const uint32_t sb = i0 * 64u + (uint32_t)(__ffsll((unsigned long long)m) - 1);
		const uint64_t s = ld(state(H, sb));
		if (st_count(s) != 0 || !(s & ST_LISTED))
			continue;
		f_and(w, ~((uint64_t)1 << (sb & 63u)));
		try_release(H, sb, f_and(state(H, sb), ~ST_LISTED) & ~ST_LISTED, x, x * (H->nctx / KMA_NXCD));
	}
}

/* Quiescent heap only: summary levels 1 and 2 of (x, c)'s index exactly
 * from level 0. */
__device__ inline void pidx_rebuild(const kma_dheap *H, uint32_t x, uint32_t c)
{
	uint64_t *P = pidx(H, x, c);
	const uint32_t w0 = pidx_w0(H), w1 = pidx_w1(H), w2 = (w1 + 63u) >> 6;
	for (uint32_t i1 = 0; i1 < w1; i1++) {
		uint64_t v = 0;
		for (uint32_t j = 0; j < 64u && i1 * 64u + j < w0; j++)
			v |= (uint64_t)(ld(P + i1 * 64u + j) != 0) << j;
		st(P + w0 + i1, v);
	}
	for (uint32_t i2 = 0; i2 < w2; i2++) {
		uint64_t v = 0;
		for (uint32_t j = 0; j < 64u && i2 * 64u + j < w1; j++)
			v |= (uint64_t)(ld(P + w0 + i2 * 64u + j) != 0) << j;
		st(P + w0 + w1 + i2, v);
	}
}

/* Attach listed superblock sb (just taken out of the index) to a head of
 * class c: owned, not listed, count + n, generation bumped (reservations
 * still in flight against the head it left then see a new incarnation).
 * Returns the state before. */
__device__ __forceinline__ uint64_t attach_listed(const kma_dheap *H, uint32_t sb, uint32_t n)
{
	uint64_t s = ld(state(H, sb));
	for (;;) {
		const uint64_t g = (uint64_t)((st_gen(s) + 1u) & ST_GEN_MASK) << ST_GEN_SHIFT;
		const uint64_t nv = (((s & ~(ST_LISTED | (~0ull << ST_GEN_SHIFT))) | ST_OWNED) + n) | g;
		const uint64_t seen = cas(state(H, sb), s, nv);
		if (seen == s)
			return s;
		s = seen;
	}
}

struct resv {
	uint32_t sb;
	uint32_t base;
	uint32_t got;
	uint32_t ok; /* 0: pool exhausted */
};

/* Reserve up to n slots of class c from context ctx's current superblock
 * (one lane executes this for its wavefront or work-group group).  A head
 * that another lane is refilling (HEAD_BUSY) is not waited for: the call
 * returns {ok 1, got 0} and the caller's wavefront-uniform loop retries.
 * A lane spinning here on a lock held by a lane of its own wavefront could
 * keep that lane masked off until the watchdog, and with every class of a
 * wavefront served at once, a wavefront holding one head while a lane of
 * it waits on another can close a cycle with a second wavefront. */
/* `hint`, if it names a superblock, is taken as the head's value without
 * reading the head (a value the caller saw earlier: it is checked like a
 * stale read, by the generation and class the fadd returns); *used gets the
 * head value the returned slots came from (0 when none). */
__device__ __forceinline__ resv reserve(const kma_dheap *H, uint32_t x, uint32_t ctx, uint32_t c, uint32_t n,
					uint64_t hint = 0, uint64_t *used = nullptr)
{
	uint64_t *hp = head(H, ctx, c);
	uint32_t cap = H->class_cap[c];
	waiter wt;
	uint64_t h = hint > HEAD_BUSY && (uint32_t)hint - 2u < H->nsb ? hint : ld(hp);
	if (used)
		*used = 0;
	for (;;) {
		if (wt.expired()) {
			bump_stat(H, KMA_STAT_WAIT_RESERVE);
			return resv{0, 0, 0, 0};
		}
		if (h == HEAD_EMPTY) {
			uint64_t seen = cas(hp, HEAD_EMPTY, HEAD_BUSY);
			if (seen != HEAD_EMPTY) {
				h = seen;
				continue;
			}
			/* an empty superblock from the pool search, else a listed
			 * partly free one of this XCD and class; while the XCD is under
			 * pressure (its last search found nothing) the listed ones come
			 * first, the search when none is left, and the index once more
			 * if the search found nothing either */
			uint32_t sb, gen;
			uint64_t old;
			const bool press = under_pressure(H, x);
			bool listed = press && pidx_take(H, x, c, ctx, &sb);
			if (!listed && pool_pop(H, x, ctx, 0u, &sb)) {
				old = fadd(state(H, sb), ((uint64_t)(c + 1u) << ST_CLASS_SHIFT) | ST_OWNED |
								 ((uint64_t)x << ST_HOME_SHIFT) | n);
				gen = st_gen(old);
			} else if (listed || pidx_take(H, x, c, ctx, &sb)) {
				old = attach_listed(H, sb, n);
				gen = (st_gen(old) + 1u) & ST_GEN_MASK;
				bump_stat(H, KMA_STAT_REATTACHED);
			} else {
				st(hp, HEAD_EMPTY);
				bump_stat(H, KMA_STAT_POOL_EMPTY);
				return resv{0, 0, 0, 0};
			}
			uint32_t cnt = st_count(old);
			uint32_t got = min(n, sat_sub(cap, cnt));
			if (got < n)
				fsub(state(H, sb), (uint64_t)(n - got));
			h = ((uint64_t)gen << 32) | (uint64_t)(sb + 2u);
			st(hp, h);
			if (got == 0)
				continue;
			if (used)
				*used = h;
			return resv{sb, cnt, got, 1};
		}
		if (h == HEAD_BUSY) {
			__builtin_amdgcn_s_sleep(2);
			return resv{0, 0, 0, 1};
		}
		uint32_t sb = (uint32_t)h - 2u;
		uint32_t g = (uint32_t)(h >> 32);
		uint64_t old = fadd(state(H, sb), (uint64_t)n);
		/* the incarnation read from the head: same generation and class */
		if (st_gen(old) != g || st_cls1(old) != c + 1u) {
			undo(H, sb, n, x, ctx);
			h = fadd(hp, (uint64_t)0);
			continue;
		}
		uint32_t cnt = st_count(old);
		uint32_t got = min(n, sat_sub(cap, cnt));
		if (got < n) {
			undo(H, sb, n - got, x, ctx);
			uint64_t seen = cas(hp, h, HEAD_EMPTY);
			if (seen == h) {
				after_detach(H, sb, f_and(state(H, sb), ~ST_OWNED) & ~ST_OWNED, x, ctx);
				seen = HEAD_EMPTY;
			}
			if (got == 0) {
				h = seen;
				continue;
			}
		}
		if (used)
			*used = h;
		return resv{sb, cnt, got, 1};
	}
}

/* Bits [lo, hi) that fall in bitmap word wi. */
__device__ __forceinline__ uint32_t range_bits(uint32_t wi, uint32_t lo, uint32_t hi)
{
	uint32_t a = max(lo, wi * 32u), b = min(hi, wi * 32u + 32u);
	if (a >= b)
		return 0u;
	uint32_t n = b - a;
	return (n == 32u ? ~0u : ((1u << n) - 1u)) << (a - wi * 32u);
}

/* Lane of the k-th (from 0) set bit of m; m has more than k bits set. */
__device__ __forceinline__ uint32_t nth_lane(uint64_t m, uint32_t k)
{
	uint32_t pos = 0;
	for (uint32_t s = 32; s; s >>= 1) {
		uint32_t n = (uint32_t)__popcll((m >> pos) & ((1ull << s) - 1ull));
		if (n <= k) {
			k -= n;
			pos += s;
		}
	}
	return pos;
}

/* Lowest-but-k free bit: the position of the k-th (from 0) set bit of w;
 * w has more than k bits set. */
__device__ __forceinline__ uint32_t nth_bit(uint32_t w, uint32_t k)
{
	uint32_t pos = 0;
	for (uint32_t s = 16; s; s >>= 1) {
		uint32_t n = (uint32_t)__popc((w >> pos) & ((1u << s) - 1u));
		if (n <= k) {
			k -= n;
			pos += s;
		}
	}
	return pos;
}

/* Scan superblock sb (class c) for a free bit, from word `wi` whose value
 * `w` the caller has just read.  In each word the lane takes the
 * (skip mod n)-th of its n free bits, so lanes that scan together (skip =
 * their rank) spread over different bits instead of all trying the lowest.
 * Words after the first are read with plain atomic loads; after a wrap
 * they are re-read with RMWs, so a stale cached word cannot hide free bits
 * for long.  NULL (counted) when the watchdog expires.  A caller that holds
 * a reservation always has a free bit to find: a block's bit is set only
 * under a reservation and cleared before its count is released, so the
 * state's count never falls below the set bits and the free bits are at
 * least the unserved reservations; every lost race is another reservation
 * served.  The watchdog therefore expires only on a heap whose bitmap was
 * written outside the allocator, and the slot stays counted (the heap
 * walker reports the superblock's count/bitmap mismatch). */
__device__ inline char *claim_scan(const kma_dheap *H, uint32_t sb, uint32_t c, uint32_t wi, uint32_t w,
				   uint32_t skip)
{
	uint32_t cap = H->class_cap[c];
	uint32_t *bm = bitmap(H, sb);
	uint32_t lastw = (cap - 1u) >> 5;
	uint32_t tail = (cap & 31u) ? ((1u << (cap & 31u)) - 1u) : ~0u;
	bool fresh = false;
	waiter wt;
	while (!wt.expired()) {
		uint32_t avail = ~w & (wi == lastw ? tail : ~0u);
		if (avail) {
			uint32_t bit = nth_bit(avail, skip % (uint32_t)__popc(avail));
			uint32_t m = 1u << bit;
			uint32_t old = f_or(bm + wi, m);
			if (!(old & m))
				return sbdata(H, sb) + (uint64_t)(wi * 32u + bit) * H->class_size[c];
			w = old | m;
			continue;
		}
		if (++wi > lastw) {
			wi = 0;
			fresh = true;
		}
		w = fresh ? f_or(bm + wi, 0u) : ld(bm + wi);
	}
	bump_stat(H, KMA_STAT_WAIT_CLAIM);
	return nullptr;
}

/* Claim one free bit of superblock sb, trying the reserved slot `start`
 * itself first with no read (blocks are claimed bottom-up and mostly freed
 * in bulk, so it is usually free), then scanning from its word. */
__device__ inline char *claim(const kma_dheap *H, uint32_t sb, uint32_t c, uint32_t start, uint32_t skip = 0u)
{
	uint32_t cap = H->class_cap[c];
	uint32_t b = start < cap ? start : 0u;
	uint32_t m = 1u << (b & 31u);
	uint32_t w = f_or(bitmap(H, sb) + (b >> 5), m);
	if (!(w & m))
		return sbdata(H, sb) + (uint64_t)b * H->class_size[c];
	return claim_scan(H, sb, c, b >> 5, w | m, skip);
}

/* ---- large requests: page runs ----
 * A request of k pages (1..64) takes k pages at the start of a free aligned
 * window of 2^l >= k pages inside one 64-page bitmap word (level l).  For
 * each level, psum(H, l) holds one bit per bitmap word: set when that word
 * has no free window of that level.  The bits are hints that are exact
 * whenever no call is in flight:
 *   - a caller that saw a word without a window sets the bit, waits for the
 *     set to complete, re-reads the word with an RMW and clears the bit
 *     again if a window has appeared;
 *   - a free (or the rollback of a lost claim) clears the bit of every level
 *     that gained a window, after its own bitmap update completed.
 * Whichever of a racing set and clear comes last, one of the two sides sees
 * the other's bitmap update, so a window is never hidden once both finish.
 * A second level, psum2(H, l), holds one bit per summary word, set when all
 * of its bits are, with the same protocol one level up (a setter that made
 * a summary word all ones sets its bit, waits, re-reads the summary word and
 * clears the bit again if it is no longer full; a clear that took a summary
 * word out of all ones clears its bit).  An allocation reads the top-level
 * words, then only summary words that may have a candidate, then only
 * bitmap words that may hold a window: a full page region fails after
 * psum2_words() reads. */
__device__ __forceinline__ uint64_t page_valid(const kma_dheap *H, uint64_t wi)
{
	uint64_t n = H->npages - (wi << 6);
	return n >= 64 ? ~0ull : ((1ull << n) - 1ull);
}

/* Bit j set: pages j .. j + 2^l - 1 are free and valid, j a multiple of 2^l. */
__device__ __forceinline__ uint64_t page_windows(uint64_t w, uint32_t l, uint64_t valid)
{
	uint64_t z = ~w & valid;
	for (uint32_t sh = 1; sh < (1u << l); sh <<= 1)
		z &= z >> sh;
	/* window starts: every multiple of 2^l, by doubling */
	uint64_t pat = 1;
	for (uint32_t sh = 1u << l; sh < 64u; sh <<= 1)
		pat |= pat << sh;
	return z & pat;
}

/* Summary word si of level l went all ones (the caller's set did it): set
 * its top-level bit, wait, and clear it again if the word is no longer full. */
__device__ __forceinline__ void psum2_full(const kma_dheap *H, uint32_t l, uint64_t si)
{
	uint64_t *p2 = psum2(H, l) + (si >> 6);
	uint64_t b2 = 1ull << (si & 63);
	if (f_or(p2, b2) & b2)
		return;
	order_prior();
	if (f_or(psum(H, l) + si, (uint64_t)0) != ~0ull)
		f_and(p2, ~b2);
}

/* A summary bit of word si (level l) was cleared; `old` is the word before:
 * if it was all ones, clear its top-level bit. */
__device__ __forceinline__ void psum2_gained(const kma_dheap *H, uint32_t l, uint64_t si, uint64_t old)
{
	if (old == ~0ull)
		f_and(psum2(H, l) + (si >> 6), ~((uint64_t)1 << (si & 63)));
}

__device__ __forceinline__ void psum_mark_full(const kma_dheap *H, uint32_t l, uint64_t wi)
{
	uint64_t *sp = psum(H, l) + (wi >> 6);
	uint64_t bit = 1ull << (wi & 63);
	/* already set: its setter re-checks the word (read first: many callers
	 * find the same full word, and an RMW each would queue them) */
	if (ld(sp) & bit)
		return;
	uint64_t old = f_or(sp, bit);
	if (old & bit)
		return;
	order_prior();
	uint64_t w = f_or(pbits(H) + wi, (uint64_t)0);
	if (page_windows(w, l, page_valid(H, wi)))
		psum2_gained(H, l, wi >> 6, f_and(sp, ~bit));
	else if ((old | bit) == ~0ull)
		psum2_full(H, l, wi >> 6);
}

/* Clear the hint of every level that gained a window going from `before` to
 * `after` (the caller's bitmap update has completed). */
__device__ __forceinline__ void psum_gained(const kma_dheap *H, uint64_t wi, uint64_t before, uint64_t after)
{
	uint64_t valid = page_valid(H, wi);
	const uint64_t bit = (uint64_t)1 << (wi & 63);
	for (uint32_t l = 0; l < KMA_PAGE_LEVELS; l++)
		if (!page_windows(before, l, valid) && page_windows(after, l, valid)) {
			/* a bit read clear needs no clear: a setter whose set lands
			 * after this read re-reads the word after this caller's
			 * bitmap update and clears it itself */
			uint64_t *sp = psum(H, l) + (wi >> 6);
			if (ld(sp) & bit)
				psum2_gained(H, l, wi >> 6, f_and(sp, ~bit));
		}
}

/* Set the hint of every level that lost its last window: the protocol of
 * psum_mark_full, with one wait and one re-read for all the levels. */
__device__ __forceinline__ void psum_lost(const kma_dheap *H, uint64_t wi, uint64_t before, uint64_t after)
{
	uint64_t valid = page_valid(H, wi);
	uint64_t bit = 1ull << (wi & 63);
	uint32_t set = 0, full = 0;
	for (uint32_t l = 0; l < KMA_PAGE_LEVELS; l++) {
		if (!page_windows(before, l, valid) || page_windows(after, l, valid))
			continue;
		if (ld(psum(H, l) + (wi >> 6)) & bit)
			continue; /* already set: its setter re-checks the word */
		uint64_t old = f_or(psum(H, l) + (wi >> 6), bit);
		if (!(old & bit)) {
			set |= 1u << l;
			full |= (old | bit) == ~0ull ? 1u << l : 0u;
		}
	}
	if (!set)
		return;
	order_prior();
	uint64_t w = f_or(pbits(H) + wi, (uint64_t)0);
	for (uint32_t l = 0; l < KMA_PAGE_LEVELS; l++) {
		if (!((set >> l) & 1u))
			continue;
		if (page_windows(w, l, valid))
			psum2_gained(H, l, wi >> 6, f_and(psum(H, l) + (wi >> 6), ~bit));
		else if ((full >> l) & 1u)
			psum2_full(H, l, wi >> 6);
	}
}

/* A request examines at most this many candidate bitmap words (words the
 * summaries say may hold a window) before it gives up.  On a nearly full
 * page region the few windows left draw every caller at once; beyond the
 * first few candidates a caller mostly finds words that others just
 * filled, and returns NULL after a bounded number of round trips instead of
 * walking every summary word. */
#ifndef KMA_PAGE_TRIES
#define KMA_PAGE_TRIES 2u
#endif
__device__ __forceinline__ void *page_alloc(const kma_dheap *H, uint64_t size, uint32_t seed)
{
	uint64_t psz = 1ull << H->page_log2;
	uint64_t k = (size + psz - 1) >> H->page_log2;
	if (k > 64 || H->npages == 0)
		return nullptr;
	uint32_t l = 0;
	while ((1ull << l) < k)
		l++;
	uint64_t run = (k == 64) ? ~0ull : ((1ull << k) - 1ull);
	const uint64_t ns = psum_words(H), ns2 = psum2_words(H);
	uint64_t *sp = psum(H, l), *sp2 = psum2(H, l);
	uint64_t *pb = pbits(H);
	uint32_t h = seed * 0x9E3779B1u;
	const uint64_t si0 = (uint64_t)(h >> 8) % ns;
	const uint32_t rot = h & 63u, rot2 = (uint32_t)(si0 & 63u);
	uint32_t tries = 0;
	waiter wt;
	for (uint64_t jp = 0; jp < ns2; jp++) {
		uint64_t t = (si0 >> 6) + jp;
		t = t >= ns2 ? t - ns2 : t;
		uint64_t c2 = ~ld(sp2 + t);
		while (c2) {
			/* summary words from a seed-dependent one on */
			uint64_t r2 = (c2 >> rot2) | (rot2 ? c2 << (64u - rot2) : 0ull);
			uint32_t b2 = ((uint32_t)__ffsll((unsigned long long)r2) - 1u + rot2) & 63u;
			c2 &= ~(1ull << b2);
			const uint64_t si = (t << 6) + b2;
			if (si >= ns)
				continue;
			uint64_t cand = ~ld(sp + si);
			while (cand) {
				if (wt.expired()) {
					bump_stat(H, KMA_STAT_WAIT_POOL);
					return nullptr;
				}
				/* start at a seed-dependent bit so callers spread out */
				uint64_t r = (cand >> rot) | (rot ? cand << (64u - rot) : 0ull);
				uint32_t b = ((uint32_t)__ffsll((unsigned long long)r) - 1u + rot) & 63u;
				uint64_t wi = (si << 6) + b;
				if (wi >= ((H->npages + 63) >> 6)) { /* no such word */
					cand &= ~(1ull << b);
					continue;
				}
				if (++tries > KMA_PAGE_TRIES)
					return nullptr;
				uint64_t valid = page_valid(H, wi);
				uint64_t w = ld(pb + wi);
				for (;;) {
					if (wt.expired()) {
						bump_stat(H, KMA_STAT_WAIT_POOL);
						return nullptr;
					}
					uint64_t win = page_windows(w, l, valid);
					if (!win) {
						psum_mark_full(H, l, wi);
						cand &= ~(1ull << b);
						break;
					}
					uint32_t j = (uint32_t)__ffsll((unsigned long long)win) - 1u;
					uint64_t m = run << j;
					uint64_t old = f_or(pb + wi, m);
					if ((old & m) == 0) {
						uint64_t pi = (wi << 6) + j;
						st(plen(H) + pi, (uint8_t)k);
						psum_lost(H, wi, old, old | m);
						return base(H) + H->off_pages + (pi << H->page_log2);
					}
					/* lost a race: give back the pages this call took */
					uint64_t mine = m & ~old;
					if (mine) {
						uint64_t o2 = f_and(pb + wi, ~mine);
						order_prior();
						psum_gained(H, wi, o2, o2 & ~mine);
						w = o2 & ~mine;
					} else {
						w = old;
					}
				}
			}
		}
	}
	return nullptr;
}

__device__ inline void page_free(const kma_dheap *H, char *p)
{
	uint64_t pi = (uint64_t)(p - (base(H) + H->off_pages)) >> H->page_log2;
	if (pi >= H->npages) {
		bump_stat(H, KMA_STAT_BAD_FREE);
		return;
	}
	uint32_t k = ld(plen(H) + pi);
	if (k == 0 || k > 64) {
		bump_stat(H, KMA_STAT_BAD_FREE);
		return;
	}
	uint64_t wi = pi >> 6;
	uint64_t m = ((k == 64) ? ~0ull : ((1ull << k) - 1ull)) << (pi & 63);
	/* pages are reused by any XCD: write back this XCD's L2 first */
	release_l2();
	uint64_t old = f_and(pbits(H) + wi, ~m);
	if ((old & m) != m)
		bump_stat(H, KMA_STAT_DOUBLE_FREE);
	order_prior();
	psum_gained(H, wi, old, old & ~m);
}

/* After a kernel boundary every L2 has been written back, so superblocks in
 * dirty shards and context spare lines may go to any XCD: move them to
 * their XCD's clean shards and drop pending clean requests (one thread; no
 * allocation may be in flight). */
__device__ inline void clean_all(const kma_dheap *H)
{
	uint32_t per = H->nctx / KMA_NXCD;
	for (uint32_t x = 0; x < KMA_NXCD; x++) {
		st(creq(H, x), (uint64_t)0);
		uint32_t sb, n = 0;
		for (uint32_t k = 0; k < dirty_per_xcd(H); k++)
			while (pool_pop_shard(H, dirty_shard(H, x, k), &sb))
				pool_push(H, clean_shard(H, x, n++), sb);
		for (uint32_t c = x * per; c < (x + 1) * per; c++)
			for (uint32_t k = 0; k < H->spare_slots; k++) {
				uint32_t v = xchg(spare(H, c) + k, 0u);
				if (v)
					pool_push(H, clean_shard(H, x, n++), v - 1u);
			}
	}
}

/* ---- initialisation (every metadata word starts at zero) ----
 * Item i of max(nsb, npool, KMA_PAGE_LEVELS * psum_words) items:
 *   - superblocks start in the clean shards: clean shard j (of the
 *     C = 8 * (S - D) clean shards, S - D per XCD) holds superblocks j,
 *     j + C, j + 2C, ... (lowest on top); dirty shards start empty;
 *   - each XCD's shard mask has the bits of its non-empty clean shards;
 *   - a page-window summary bit starts set only for a word that does not
 *     exist or cannot hold a window of that level (a partial last word), a
 *     top-level bit only for a summary word that does not exist or is all
 *     ones. */
__device__ __forceinline__ uint64_t init_shard_len(uint64_t j, uint64_t nsb, uint64_t C)
{
	return j < nsb ? (nsb - j + C - 1u) / C : 0u;
}

/* Summary word si of level l in the empty page region: a bit set only for a
 * bitmap word that does not exist or cannot hold a window of the level. */
__device__ inline uint64_t init_psum_word(const kma_dheap *H, uint32_t l, uint64_t si)
{
	uint64_t nwords = (H->npages + 63) >> 6, bits = 0;
	for (uint32_t b = 0; b < 64; b++) {
		uint64_t wi = (si << 6) + b;
		if (wi >= nwords || !page_windows(0, l, page_valid(H, wi)))
			bits |= (uint64_t)1 << b;
	}
	return bits;
}

__device__ inline void init_item(const kma_dheap *H, uint64_t i)
{
	uint64_t nsb = H->nsb;
	uint32_t S = shards_per_xcd(H), D = dirty_per_xcd(H), K = S - D, C = K * KMA_NXCD;
	if (i < nsb) {
		uint64_t nx = i + C;
		*pnext(H, i) = nx < nsb ? (uint32_t)(nx + 1u) : 0u;
	}
	if (i < C) {
		uint64_t *hp = pool(H, (uint32_t)(i / K) * S + D + (uint32_t)(i % K));
		hp[0] = i < nsb ? (uint64_t)(i + 1u) : 0u;
		hp[1] = init_shard_len(i, nsb, C);
	}
	if (i < KMA_NXCD) {
		uint64_t m = 0;
		for (uint64_t j = i * K; j < (i + 1) * K; j++)
			if (init_shard_len(j, nsb, C))
				m |= (uint64_t)1 << (D + (j - i * K));
		*pmask(H, (uint32_t)i) = m;
	}
	uint64_t ns = psum_words(H), ns2 = psum2_words(H);
	if (H->npages && i < KMA_PAGE_LEVELS * ns) {
		uint32_t l = (uint32_t)(i / ns);
		uint64_t si = i % ns;
		psum(H, l)[si] = init_psum_word(H, l, si);
	}
	if (H->npages && i < KMA_PAGE_LEVELS * ns2) {
		uint32_t l = (uint32_t)(i / ns2);
		uint64_t j = i % ns2, bits = 0;
		for (uint32_t b = 0; b < 64; b++) {
			uint64_t si = (j << 6) + b;
			if (si >= ns || init_psum_word(H, l, si) == ~0ull)
				bits |= (uint64_t)1 << b;
		}
		psum2(H, l)[j] = bits;
	}
}

__host__ __device__ inline uint64_t init_items(uint64_t nsb, uint32_t npool, uint64_t npages)
{
	uint64_t n = nsb > npool ? nsb : npool;
	uint64_t nps = KMA_PAGE_LEVELS * ((((npages + 63) >> 6) + 63) >> 6);
	return nps > n ? nps : n;
}

} /* namespace kma_nat */

/* Superblocks currently free for any class, counted over slice `part` of
 * `nparts` of the heap's free lists (shard counts, then context spare
 * slots) by a full wavefront (all 64 lanes call it; every lane gets the
 * slice's sum, modulo 2^64).  Summing the slices of parts 0 .. nparts-1,
 * computed by as many wavefronts, gives the total at a fraction of one
 * wavefront's time; a pop's decrement of a shard count can land before its
 * push's increment, so read the total as signed and clamp it at zero.  A
 * snapshot: pushes and pops in flight may be missed. */
__device__ inline uint64_t kma_free_superblocks_part(const kma_dheap *H, uint32_t part, uint32_t nparts)
{
	using namespace kma_nat;
	const uint32_t n = H->spare_slots, wpc = (n + 1u) / 2u; /* spare words per line */
	const uint64_t ns = H->npool, N = ns + (uint64_t)H->nctx * wpc;
	const uint64_t lo = N * part / nparts, hi = N * (part + 1u) / nparts;
	uint64_t s = 0;
	for (uint64_t i = lo + lane(); i < hi; i += 4u * 64u) {
		uint64_t v[4];
#pragma unroll
		for (uint32_t u = 0; u < 4u; u++) {
			uint64_t j = i + u * 64u;
			if (j >= hi)
				v[u] = 0u;
			else if (j < ns)
				v[u] = ld(pool(H, (uint32_t)j) + 1);
			else
				v[u] = ld((const uint64_t *)spare(H, (uint32_t)((j - ns) / wpc)) + (j - ns) % wpc);
		}
#pragma unroll
		for (uint32_t u = 0; u < 4u; u++) {
			uint64_t j = i + u * 64u;
			if (j >= hi)
				continue;
			if (j < ns) {
				s += v[u];
			} else {
				uint32_t w = (uint32_t)((j - ns) % wpc);
				s += ((uint32_t)v[u] != 0u ? 1u : 0u) +
				     (2u * w + 1u < n && (uint32_t)(v[u] >> 32) != 0u ? 1u : 0u);
			}
		}
	}
	for (int o = 32; o > 0; o >>= 1)
		s += __shfl_xor(s, o);
	return s;
}

/* Superblocks currently free for any class (all of the heap's free lists,
 * by one full wavefront; see kma_free_superblocks_part). */
__device__ inline uint64_t kma_pooled_superblocks(const kma_dheap *H)
{
	uint64_t s = kma_free_superblocks_part(H, 0u, 1u);
	return (int64_t)s < 0 ? 0u : s;
}

/* Allocate `size` bytes; NULL for size 0 or when the heap cannot serve it.
 * The calling lanes of a wavefront are served together: grouped by size
 * class with ballots, the first lane of each class reserves slots for its
 * class from the context's head (every class at once), then the first lane
 * in each bitmap word of a class's reserved slots sets that word's bits with
 * one RMW for all of them; a lane whose slot was already taken scans for
 * another free bit.  Lanes left without a slot (a busy head, a superblock
 * that filled up) go round again.  Lanes with no small request (size 0, or a
 * page run served on their own) stay with the wavefront until it leaves the
 * loop together. */
__device__ __forceinline__ void *kma_malloc(kma_dheap *H, size_t size)
{
	using namespace kma_nat;
	const bool small = size != 0 && size <= H->max_small;
	char *result = nullptr;
	if (size > H->max_small) {
		uint32_t seed = (blockIdx.x * blockDim.x + threadIdx.x) ^ (uint32_t)clock64();
		result = (char *)page_alloc(H, size, seed);
	}
	if (!__ballot(small))
		return result;
	const uint32_t c = small ? (uint32_t)class_of(H, (uint32_t)size) : 0u;
	const uint32_t x = xcc();
	const uint32_t ctx = ctx_of(H, x);
	bool pending = small;
	waiter wt;
	for (;;) {
		uint64_t left = __ballot(pending);
		if (left == 0)
			break;
		if (__ballot(wt.expired())) {
			/* a head stayed busy for the whole watchdog period */
			if (lane() == (uint32_t)__ffsll((unsigned long long)left) - 1u)
				fadd(stats(H) + KMA_STAT_WAIT_RESERVE, (uint64_t)__popcll(left));
			break;
		}
		/* class groups of the pending lanes (no memory traffic) */
		uint64_t grp = 0;
		do {
			uint32_t l0 = (uint32_t)__ffsll((unsigned long long)left) - 1u;
			uint32_t lc = rdlane(c, l0);
			uint64_t g = __ballot(pending && c == lc);
			if (pending && c == lc)
				grp = g;
			left &= ~g;
		} while (left);
		const uint32_t q = (uint32_t)__popcll(grp & lanemask_lt());
		resv r = {0, 0, 0, 0};
		if (pending && q == 0u)
			r = reserve(H, x, ctx, c, (uint32_t)__popcll(grp));
		const int src = (int)(pending ? (uint32_t)__ffsll((unsigned long long)grp) - 1u : lane());
		const uint32_t sb = __shfl(r.sb, src), rbase = __shfl(r.base, src);
		const uint32_t got = __shfl(r.got, src), ok = __shfl(r.ok, src);
		const bool mine = pending && ok && q < got;
		if (pending && !ok)
			pending = false;
		const uint32_t slot = rbase + q, wi = slot >> 5;
		uint32_t old = 0;
		if (mine && (q == 0u || (slot & 31u) == 0u))
			old = f_or(bitmap(H, sb) + wi, range_bits(wi, rbase, rbase + got));
		const uint32_t first = (wi << 5) > rbase ? (wi << 5) - rbase : 0u;
		old = __shfl(old, (int)(mine ? nth_lane(grp, first) : lane()));
		/* lanes whose slot was taken scan on from the word's new value,
		 * spread over its free bits by their rank among such lanes */
		const uint32_t m = 1u << (slot & 31u);
		const bool lost = mine && (old & m);
		const uint32_t rank = (uint32_t)__popcll(__ballot(lost) & lanemask_lt());
		if (mine) {
			result = !lost ? sbdata(H, sb) + (uint64_t)slot * H->class_size[c]
				       : claim_scan(H, sb, c, wi, old | range_bits(wi, rbase, rbase + got), rank);
			pending = false;
		}
	}
	return result;
}

/* Free a block obtained from kma_malloc on the same heap; NULL is a no-op.
 * A second free of the same block is detected and counted, not applied.
 * Lanes with no superblock block (NULL, a page run) stay with the
 * wavefront until it leaves together. */
__device__ __forceinline__ void kma_free(kma_dheap *H, void *ptr)
{
	using namespace kma_nat;
	char *p = (char *)ptr;
	char *d0 = base(H) + H->off_data;
	uint64_t off = (uint64_t)(p - d0);
	const bool small = p && p >= d0 && off < (H->nsb << H->sb_log2);
	if (p && !small) {
		if (H->npages && p >= base(H) + H->off_pages)
			page_free(H, p);
		else
			bump_stat(H, KMA_STAT_BAD_FREE);
	}
	if (!__ballot(small))
		return;
	uint32_t x = xcc();
	uint32_t sb = small ? (uint32_t)(off >> H->sb_log2) : 0u;
	uint64_t s = small ? ld(state(H, sb)) : 0u;
	uint32_t c1 = st_cls1(s);
	bool live = c1 != 0;
	/* the caller's stores to the block complete before its bit is
	 * released; a block homed on another XCD also needs this XCD's L2
	 * written back (one write-back and one counter update per wavefront) */
	uint64_t rm = __ballot(live && st_home(s) != x);
	if (rm) {
		if (lane() == (uint32_t)__ffsll((unsigned long long)rm) - 1u)
			fadd(stats(H) + KMA_STAT_REMOTE_FREE, (uint64_t)__popcll(rm));
		release_l2();
	} else {
		order_prior();
	}
	if (small && !live)
		bump_stat(H, KMA_STAT_BAD_FREE);
	uint32_t idx = 0;
	if (live) {
		uint64_t inoff = off & ((1ull << H->sb_log2) - 1ull);
		idx = (uint32_t)((inoff * H->class_magic[c1 - 1u]) >> 40);
	}
	/* superblock groups of the lanes with a block (no memory traffic) */
	uint64_t left = __ballot(live), grp = 0;
	while (left) {
		uint32_t l0 = (uint32_t)__ffsll((unsigned long long)left) - 1u;
		uint32_t lsb = rdlane(sb, l0);
		uint64_t g = __ballot(live && sb == lsb);
		if (live && sb == lsb)
			grp = g;
		left &= ~g;
	}
	/* bit release, one RMW per run of a group's members (in lane order)
	 * freeing consecutive slots of one bitmap word, as blocks claimed
	 * together come back: the run's last member clears the run's bits, its
	 * first member found from the ballot of run starts; every lane reads
	 * its own bit in the word the RMW returned, so a double free is still
	 * seen by the lane that made it */
	const uint32_t l = lane();
	const uint64_t le = (2ull << l) - 1ull; /* lanes 0..l */
	const uint64_t below = grp & lanemask_lt();
	const uint32_t pidx = __shfl(idx, (int)(below ? 63u - (uint32_t)__clzll((unsigned long long)below) : l));
	const bool cont = live && below != 0u && pidx + 1u == idx && (idx & 31u) != 0u;
	const uint64_t heads = __ballot(live && !cont), conts = __ballot(cont);
	const uint64_t above = grp & ~le;
	const bool last = live && !(above && ((conts >> (__ffsll((unsigned long long)above) - 1)) & 1u));
	const uint64_t lasts = __ballot(last);
	uint32_t word = 0;
	if (last) {
		const uint32_t first = 63u - (uint32_t)__clzll((unsigned long long)(heads & grp & le));
		const uint32_t span = (uint32_t)__popcll(grp & le & ~((1ull << first) - 1ull)) - 1u;
		const uint32_t hi = idx & 31u, lo = hi - span, n = span + 1u;
		const uint32_t mask = (n == 32u ? ~0u : ((1u << n) - 1u)) << lo;
		word = f_and(bitmap(H, sb) + (idx >> 5), ~mask);
	}
	const uint32_t src =
		live ? (uint32_t)__ffsll((unsigned long long)(lasts & grp & ~lanemask_lt())) - 1u : l;
	word = __shfl(word, (int)src);
	if (live && !(word & (1u << (idx & 31u)))) {
		bump_stat(H, KMA_STAT_DOUBLE_FREE);
		live = false;
	}
	/* one count update per superblock per wavefront (its live members),
	 * every superblock at once */
	const uint64_t lv = __ballot(live);
	uint32_t ctx = ctx_of(H, x);
	if (live && l == (uint32_t)__ffsll((unsigned long long)(grp & lv)) - 1u) {
		uint32_t n = (uint32_t)__popcll(grp & lv);
		after_sub(H, sb, fsub(state(H, sb), (uint64_t)n) - n, n, x, ctx);
	}
}

#endif /* KMA_DEVICE_H */
