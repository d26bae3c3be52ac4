/*
 * kma_detbatch.hip - deterministic batch mode on the reference layout
 * (kma_batch_malloc_det / kma_batch_free_det in include/kma_work.h).
 *
 * A batch of n mallocs is served in parallel with exactly the result of
 * running malloc(sizes[0]) .. malloc(sizes[n-1]) one at a time in index
 * order (kma.cl:279-303), starting from a heap with no live block and no
 * attached class superblock.  In that serialized order:
 *   - request i of class c is the k-th request of its class (k = its rank
 *     among class-c requests by index); group q = k / N_c of the class is
 *     served by one superblock, whose refill happens at the group's first
 *     request (k = q N_c);
 *   - refills pop the free queue in index order of those first requests,
 *     and once a pop fails every later one fails too (no frees in the
 *     batch): group q gets the g-th popped superblock, g = the global rank
 *     of its first request among all first requests, iff g < P (P pops
 *     succeed);
 *   - inside a fresh superblock the r-th request (r = k mod N_c) decrements
 *     the free count to N_c - 1 - r and claims block N_c - 1 - r, writing
 *     that ticket into the block (kma.cl:180-192, 232-266).
 * Ranks come from per-block class counts and prefix sums; only the P pops
 * run on one lane, with the same queue code as the allocator, so the queue
 * words and their tags are the serialized ones.
 *
 * kma_batch_free_det frees that batch's blocks as free(out[0]) ..
 * free(out[n-1]) would (kma.cl:305-373): block b of superblock s writes the
 * bitmap word as it stood before its own bit was cleared, i.e. the bits of
 * the blocks of that word from the lowest block the batch allocated in s up
 * to b (the ones freed later are the lower ones); a superblock empties, is
 * unlinked and enqueued at the free of its lowest block, so superblocks are
 * enqueued in index order of those frees, again on one lane.
 */
#include <hip/hip_runtime.h>
#include <pthread.h>

#include "kma_launch.h"
#include "kma_ref_device.h"
#include "kma_work.h"

#define KMA_CHECK(x)                                   \
	do {                                           \
		hipError_t e_ = (x);                   \
		if (e_ != hipSuccess)                  \
			return -(int)e_;               \
	} while (0)

#define DB 256u   /* threads per block */
#define DCLS 16u  /* classes (log2 <= 17) */

using namespace kma_ref;

namespace {

__device__ __forceinline__ uint32_t class_slots(const view &v, int c)
{
	return slots_by_size(v.log2, size_by_sbid(v.log2, c));
}

/* Exclusive scan of one value per thread over the block (DB threads). */
__device__ uint32_t block_excl_scan(uint32_t x, uint32_t *tmp, uint32_t *total)
{
	uint32_t t = threadIdx.x;
	tmp[t] = x;
	__syncthreads();
	for (uint32_t o = 1; o < DB; o <<= 1) {
		uint32_t y = t >= o ? tmp[t - o] : 0u;
		__syncthreads();
		tmp[t] += y;
		__syncthreads();
	}
	*total = tmp[DB - 1];
	uint32_t r = tmp[t] - x;
	__syncthreads();
	return r;
}

} // namespace

/* class of each request; per-block class counts; requests of a 1-slot
 * class are counted in *err (serialized, the reference blocks on the second
 * one, kma.cl:203-204, so the batch is refused) */
__global__ void __launch_bounds__(DB) k_det_class(clheap *heap, const uint32_t *sizes, uint64_t n, int8_t *cls,
						  uint32_t *cnt, uint32_t *err)
{
	view v = make_view(heap);
	__shared__ uint32_t cs[DCLS];
	uint32_t t = threadIdx.x;
	if (t < DCLS)
		cs[t] = 0;
	__syncthreads();
	uint64_t i = (uint64_t)blockIdx.x * DB + t;
	int c = -1;
	if (i < n) {
		c = sbid_by_size(v.log2, sizes[i]);
		cls[i] = (int8_t)c;
	}
	if (c >= 0 && class_slots(v, c) < 2)
		atomicAdd(err, 1u);
	if (c >= 0)
		atomicAdd(&cs[c], 1u);
	__syncthreads();
	if (t < DCLS)
		cnt[(uint64_t)blockIdx.x * DCLS + t] = cs[t];
}

/* column `blockIdx.x` of a [nb][ncol] count table -> exclusive prefix over
 * rows, column total in tot[col] */
__global__ void __launch_bounds__(DB) k_det_scan(uint32_t *tab, uint64_t nb, uint32_t ncol, uint32_t *tot)
{
	__shared__ uint32_t tmp[DB];
	uint32_t col = blockIdx.x, carry = 0;
	for (uint64_t base = 0; base < nb; base += DB) {
		uint64_t r = base + threadIdx.x;
		uint32_t x = r < nb ? tab[r * ncol + col] : 0u;
		uint32_t total;
		uint32_t e = block_excl_scan(x, tmp, &total);
		if (r < nb)
			tab[r * ncol + col] = carry + e;
		carry += total;
	}
	if (threadIdx.x == 0)
		tot[col] = carry;
}

/* stable rank within the class; refill-event flags and per-block counts */
__global__ void __launch_bounds__(DB) k_det_rank(clheap *heap, const int8_t *cls, uint64_t n, const uint32_t *off,
						 uint32_t *rank, uint32_t *evcnt)
{
	view v = make_view(heap);
	__shared__ uint32_t wc[DB / 64][DCLS];
	__shared__ uint32_t ev[DB / 64];
	uint32_t t = threadIdx.x, w = t >> 6, l = t & 63u;
	uint64_t i = (uint64_t)blockIdx.x * DB + t;
	int c = i < n ? cls[i] : -1;
	uint32_t wr = 0;
	for (uint32_t k = 0; k < DCLS; k++) {
		uint64_t m = __ballot(c == (int)k);
		if (c == (int)k)
			wr = (uint32_t)__popcll(m & ((1ull << l) - 1ull));
		if (l == 0)
			wc[w][k] = (uint32_t)__popcll(m);
	}
	__syncthreads();
	bool e = false;
	if (c >= 0) {
		uint32_t r = off[(uint64_t)blockIdx.x * DCLS + c] + wr;
		for (uint32_t w2 = 0; w2 < w; w2++)
			r += wc[w2][c];
		rank[i] = r;
		e = r % class_slots(v, c) == 0;
	}
	uint64_t m = __ballot(e);
	if (l == 0)
		ev[w] = (uint32_t)__popcll(m);
	__syncthreads();
	if (t == 0) {
		uint32_t s = 0;
		for (uint32_t k = 0; k < DB / 64; k++)
			s += ev[k];
		evcnt[blockIdx.x] = s;
	}
}

/* global rank of each refill event -> per-class group table and the
 * event's (class, group) */
__global__ void __launch_bounds__(DB) k_det_events(clheap *heap, const int8_t *cls, const uint32_t *rank, uint64_t n,
						   const uint32_t *evoff, const uint32_t *coff, uint32_t *table,
						   uint32_t *evc, uint32_t *evq)
{
	view v = make_view(heap);
	__shared__ uint32_t ev[DB / 64];
	uint32_t t = threadIdx.x, w = t >> 6, l = t & 63u;
	uint64_t i = (uint64_t)blockIdx.x * DB + t;
	int c = i < n ? cls[i] : -1;
	uint32_t N = c >= 0 ? class_slots(v, c) : 1u;
	bool e = c >= 0 && rank[i] % N == 0;
	uint64_t m = __ballot(e);
	if (l == 0)
		ev[w] = (uint32_t)__popcll(m);
	__syncthreads();
	if (e) {
		uint32_t g = evoff[blockIdx.x] + (uint32_t)__popcll(m & ((1ull << l) - 1ull));
		for (uint32_t w2 = 0; w2 < w; w2++)
			g += ev[w2];
		uint32_t q = rank[i] / N;
		table[coff[c] + q] = g;
		evc[g] = (uint32_t)c;
		evq[g] = q;
	}
}

/* first table index of each class: coff[c] = sum over c' < c of the
 * number of groups of c' */
__global__ void k_det_coff(clheap *heap, const uint32_t *tot, uint32_t *coff)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	view v = make_view(heap);
	uint32_t acc = 0;
	for (uint32_t c = 0; c < DCLS; c++) {
		coff[c] = acc;
		if (c < v.buckets) {
			uint32_t N = class_slots(v, (int)c);
			acc += (tot[c] + N - 1u) / N;
		}
	}
}

/* the refills' pops, in event order, on one lane */
__global__ void k_det_pops(clheap *heap, const uint32_t *nevp, uint32_t *popped, uint32_t *npop)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	view v = make_view(heap);
	uint32_t j = 0, nev = *nevp;
	for (; j < nev; j++) {
		char *sb = q_dequeue(v);
		if (!sb)
			break;
		popped[j] = (uint32_t)((uint64_t)(sb - v.sb0) >> v.log2);
	}
	*npop = j;
}

/* each refilled superblock's header, cleared bitmap and the class head */
__global__ void __launch_bounds__(DB) k_det_refill(clheap *heap, const uint32_t *npop, const uint32_t *popped,
						   const uint32_t *evc, const uint32_t *evq, const uint32_t *tot)
{
	view v = make_view(heap);
	uint32_t g = blockIdx.x * DB + threadIdx.x;
	if (g >= *npop)
		return;
	int c = (int)evc[g];
	uint32_t q = evq[g], size = size_by_sbid(v.log2, c), N = slots_by_size(v.log2, size);
	uint32_t left = tot[c] - q * N, used = left < N ? left : N;
	char *sb = v.sb0 + ((uint64_t)popped[g] << v.log2);
	*sb_size(sb) = size;
	*sb_state(sb) = (N << 16) | (N - used);
	for (uint32_t b = 0; b < N; b += 32)
		*sb_word(v, sb, b) = 0;
	/* the class head keeps a superblock with slots left (the last group);
	 * one that filled was unlinked by its last claim (kma.cl:216-219) */
	if (used < N)
		*cls_head(v, c) = (uint64_t)(uintptr_t)sb;
}

/* block, ticket and bit of every request; NULL results counted in *nulls
 * when it is given (one atomic per wavefront) */
__global__ void __launch_bounds__(DB) k_det_claim(clheap *heap, const int8_t *cls, const uint32_t *rank, uint64_t n,
						  const uint32_t *coff, const uint32_t *table, const uint32_t *npop,
						  const uint32_t *popped, uint64_t *out, unsigned long long *nulls)
{
	view v = make_view(heap);
	uint64_t i = (uint64_t)blockIdx.x * DB + threadIdx.x;
	if (i >= n)
		return;
	int c = cls[i];
	uint64_t res = 0;
	if (c >= 0) {
		uint32_t size = size_by_sbid(v.log2, c), N = slots_by_size(v.log2, size);
		uint32_t q = rank[i] / N, g = table[coff[c] + q];
		if (g < *npop) {
			uint32_t b = N - 1u - rank[i] % N;
			char *sb = v.sb0 + ((uint64_t)popped[g] << v.log2);
			char *p = sb + 12u + (uint64_t)b * size;
			*reinterpret_cast<uint32_t *>(p) = b;
			atomicOr(sb_word(v, sb, b), 1u << (b & 31u));
			res = (uint64_t)(p - v.h);
		}
	}
	out[i] = res;
	if (nulls) {
		/* lanes past n have returned; lane 0 of a wavefront with any
		 * active lane is active (it has the wavefront's lowest index) */
		uint64_t m = __ballot(res == 0);
		if (__lane_id() == 0 && m)
			atomicAdd(nulls, (unsigned long long)__popcll(m));
	}
}

/* zero `n` words (the allocation-free entry point clears its scratch
 * counters with a kernel, not a memset, so the sequence can be captured) */
__global__ void k_det_zero(uint32_t *p, uint32_t n, unsigned long long *q, uint32_t nq)
{
	uint32_t t = threadIdx.x;
	if (t < n)
		p[t] = 0;
	if (t < nq)
		q[t] = 0;
}

/* refused requests (1-slot classes) into the caller's status word */
__global__ void k_det_status(const uint32_t *small, unsigned long long *status)
{
	if (threadIdx.x == 0)
		status[1] = small[2];
}

/* ---- free ---- */
__global__ void __launch_bounds__(DB) k_detf_min(clheap *heap, const uint64_t *offs, uint64_t n, uint32_t *minblk)
{
	view v = make_view(heap);
	uint64_t i = (uint64_t)blockIdx.x * DB + threadIdx.x;
	if (i >= n || offs[i] == 0)
		return;
	char *p = v.h + offs[i];
	uint64_t s = (uint64_t)(p - v.sb0) >> v.log2;
	char *sb = v.sb0 + (s << v.log2);
	uint32_t b = (uint32_t)((uint64_t)(p - sb - 12) / *sb_size(sb));
	atomicMin(&minblk[s], b);
}

/* old bitmap words into the blocks; flags of superblock-emptying frees */
__global__ void __launch_bounds__(DB) k_detf_words(clheap *heap, const uint64_t *offs, uint64_t n,
						   const uint32_t *minblk, uint32_t *evcnt, uint32_t *last_sb)
{
	view v = make_view(heap);
	__shared__ uint32_t ev[DB / 64];
	uint32_t t = threadIdx.x, w = t >> 6, l = t & 63u;
	uint64_t i = (uint64_t)blockIdx.x * DB + t;
	bool e = false;
	if (i < n && offs[i] != 0) {
		char *p = v.h + offs[i];
		uint64_t s = (uint64_t)(p - v.sb0) >> v.log2;
		char *sb = v.sb0 + (s << v.log2);
		uint32_t b = (uint32_t)((uint64_t)(p - sb - 12) / *sb_size(sb));
		uint32_t lo = minblk[s], w0 = b & ~31u;
		uint32_t from = (lo > w0 ? lo : w0) - w0;
		uint64_t mask = ((2ull << (b & 31u)) - 1ull) & ~((1ull << from) - 1ull);
		*reinterpret_cast<uint32_t *>(p) = (uint32_t)mask;
		e = b == lo;
		if (e)
			last_sb[i] = (uint32_t)s;
	}
	uint64_t m = __ballot(e);
	if (l == 0)
		ev[w] = (uint32_t)__popcll(m);
	__syncthreads();
	if (t == 0) {
		uint32_t s = 0;
		for (uint32_t k = 0; k < DB / 64; k++)
			s += ev[k];
		evcnt[blockIdx.x] = s;
	}
}

/* enqueue order (index order of the emptying frees); emptied superblocks'
 * state and bitmap; class heads */
__global__ void __launch_bounds__(DB) k_detf_order(clheap *heap, const uint64_t *offs, uint64_t n,
						   const uint32_t *minblk, const uint32_t *evoff, const uint32_t *last_sb,
						   uint32_t *order)
{
	view v = make_view(heap);
	__shared__ uint32_t ev[DB / 64];
	uint32_t t = threadIdx.x, w = t >> 6, l = t & 63u;
	uint64_t i = (uint64_t)blockIdx.x * DB + t;
	bool e = false;
	uint32_t s = 0;
	if (i < n && offs[i] != 0) {
		char *p = v.h + offs[i];
		s = (uint32_t)((uint64_t)(p - v.sb0) >> v.log2);
		char *sb = v.sb0 + ((uint64_t)s << v.log2);
		uint32_t b = (uint32_t)((uint64_t)(p - sb - 12) / *sb_size(sb));
		e = b == minblk[s];
	}
	uint64_t m = __ballot(e);
	if (l == 0)
		ev[w] = (uint32_t)__popcll(m);
	__syncthreads();
	if (e) {
		uint32_t o = evoff[blockIdx.x] + (uint32_t)__popcll(m & ((1ull << l) - 1ull));
		for (uint32_t w2 = 0; w2 < w; w2++)
			o += ev[w2];
		order[o] = last_sb[i];
		char *sb = v.sb0 + ((uint64_t)s << v.log2);
		uint32_t N = *sb_state(sb) >> 16;
		*sb_state(sb) = N << 16;
		for (uint32_t b = 0; b < N; b += 32)
			*sb_word(v, sb, b) = 0;
	}
	if (blockIdx.x == 0 && t < v.buckets)
		*cls_head(v, (int)t) = 0;
}

__global__ void k_detf_enqueue(clheap *heap, const uint32_t *order, const uint32_t *nep)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	view v = make_view(heap);
	uint32_t ne = *nep;
	for (uint32_t k = 0; k < ne; k++)
		q_enqueue(v, v.sb0 + ((uint64_t)order[k] << v.log2));
}

static inline uint32_t nblk(uint64_t n)
{
	return (uint32_t)((n + DB - 1) / DB);
}

/* ---- launchers (all on the null stream, synchronous) ---- */
extern "C" int kmad_det_scan(uint32_t *tab, uint64_t nrows, uint32_t ncol, uint32_t *tot)
{
	hipLaunchKernelGGL(k_det_scan, dim3(ncol), dim3(DB), 0, 0, tab, nrows, ncol, tot);
	KMA_CHECK(hipGetLastError());
	return 0;
}

extern "C" int kmad_detf_min(void *heap, const uint64_t *offs, uint64_t n, uint32_t *minblk)
{
	hipLaunchKernelGGL(k_detf_min, dim3(nblk(n)), dim3(DB), 0, 0, (clheap *)heap, offs, n, minblk);
	KMA_CHECK(hipGetLastError());
	return 0;
}

extern "C" int kmad_detf_words(void *heap, const uint64_t *offs, uint64_t n, const uint32_t *minblk, uint32_t *evcnt,
			       uint32_t *last_sb)
{
	hipLaunchKernelGGL(k_detf_words, dim3(nblk(n)), dim3(DB), 0, 0, (clheap *)heap, offs, n, minblk, evcnt, last_sb);
	KMA_CHECK(hipGetLastError());
	return 0;
}

extern "C" int kmad_detf_order(void *heap, const uint64_t *offs, uint64_t n, const uint32_t *minblk,
			       const uint32_t *evoff, const uint32_t *last_sb, uint32_t *order)
{
	hipLaunchKernelGGL(k_detf_order, dim3(nblk(n)), dim3(DB), 0, 0, (clheap *)heap, offs, n, minblk, evoff, last_sb,
			   order);
	KMA_CHECK(hipGetLastError());
	return 0;
}

extern "C" int kmad_detf_enqueue(void *heap, const uint32_t *order, const uint32_t *ne)
{
	hipLaunchKernelGGL(k_detf_enqueue, dim3(1), dim3(64), 0, 0, (clheap *)heap, order, ne);
	KMA_CHECK(hipGetLastError());
	return 0;
}

/* The whole malloc batch on `stream`: every kernel of kma_batch_malloc_det in
 * order.  b->small[2] (refusals) must be zero beforehand; *nulls (optional)
 * gets the NULL results added. */
extern "C" int kmad_det_malloc(void *heap, const uint32_t *sizes, uint64_t n, const kmad_det_bufs *b, uint64_t *out,
			       unsigned long long *nulls, void *stream)
{
	hipStream_t s = (hipStream_t)stream;
	clheap *hp = (clheap *)heap;
	uint64_t nb = (n + DB - 1) / DB;
	uint32_t *nev = b->small, *npop = b->small + 1, *err = b->small + 2, *coff = b->small + 16;
	hipLaunchKernelGGL(k_det_class, dim3(nblk(n)), dim3(DB), 0, s, hp, sizes, n, b->cls, b->cnt, err);
	hipLaunchKernelGGL(k_det_scan, dim3(DCLS), dim3(DB), 0, s, b->cnt, nb, DCLS, b->tot);
	hipLaunchKernelGGL(k_det_rank, dim3(nblk(n)), dim3(DB), 0, s, hp, b->cls, n, b->cnt, b->rank, b->evcnt);
	hipLaunchKernelGGL(k_det_scan, dim3(1), dim3(DB), 0, s, b->evcnt, nb, 1u, nev);
	hipLaunchKernelGGL(k_det_coff, dim3(1), dim3(64), 0, s, hp, b->tot, coff);
	hipLaunchKernelGGL(k_det_events, dim3(nblk(n)), dim3(DB), 0, s, hp, b->cls, b->rank, n, b->evcnt, coff,
			   b->table, b->evc, b->evq);
	hipLaunchKernelGGL(k_det_pops, dim3(1), dim3(64), 0, s, hp, nev, b->popped, npop);
	hipLaunchKernelGGL(k_det_refill, dim3(nblk(n)), dim3(DB), 0, s, hp, npop, b->popped, b->evc, b->evq, b->tot);
	hipLaunchKernelGGL(k_det_claim, dim3(nblk(n)), dim3(DB), 0, s, hp, b->cls, b->rank, n, coff, b->table, npop,
			   b->popped, out, nulls);
	KMA_CHECK(hipGetLastError());
	return 0;
}

/* ------------------------------------------------------------------ */
/* Allocation-free entry point                                          */
/* ------------------------------------------------------------------ */
/* Scratch of kma_batch_malloc_det_noalloc, one copy per GPU, held until the
 * process ends.  Carved as (bytes, for n <= KMA_DET_NOALLOC_MAX = 2^19):
 *   cls n | rank 4n | table 4(n + DCLS) | evc 4n | evq 4n |
 *   popped 4 * 65536 (a reference-layout queue holds < 65,536 superblocks) |
 *   cnt 4 * DCLS * n/256 | evcnt 4 * n/256 | tot 4 * DCLS | small 4 * 64 */
#define NA_N ((uint64_t)KMA_DET_NOALLOC_MAX)
#define NA_NB (NA_N / DB)
#define NA_CLS 0ull
#define NA_RANK (NA_CLS + NA_N)
#define NA_TABLE (NA_RANK + 4 * NA_N)
#define NA_EVC (NA_TABLE + 4 * (NA_N + DCLS))
#define NA_EVQ (NA_EVC + 4 * NA_N)
#define NA_POPPED (NA_EVQ + 4 * NA_N)
#define NA_CNT (NA_POPPED + 4 * 65536ull)
#define NA_EVCNT (NA_CNT + 4ull * DCLS * NA_NB)
#define NA_TOT (NA_EVCNT + 4 * NA_NB)
#define NA_SMALL (NA_TOT + 4 * DCLS)
#define NA_BYTES (NA_SMALL + 4 * 64)

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

/* The device address of `scratch` on each GPU (the array's name in host
 * code is only a host-side handle; each GPU holds its own copy at its own
 * address).  Resolved once per device with that device current. */
#define NA_MAXDEV 64
static void *g_scratch_dev[NA_MAXDEV];
static pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;

extern "C" int kmad_det_scratch_setup(void **out)
{
	int dev = -1;
	KMA_CHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= NA_MAXDEV)
		return -(int)hipErrorInvalidDevice;
	pthread_mutex_lock(&g_scratch_mu);
	int rc = 0;
	if (!g_scratch_dev[dev]) {
		void *p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			rc = -(int)(e != hipSuccess ? e : hipErrorInvalidSymbol);
		else
			g_scratch_dev[dev] = p;
	}
	if (out)
		*out = g_scratch_dev[dev];
	pthread_mutex_unlock(&g_scratch_mu);
	return rc;
}

extern "C" uint64_t kmad_det_scratch_bytes(void)
{
	return NA_BYTES;
}

extern "C" int kmad_det_malloc_noalloc(void *heap, const uint32_t *sizes, uint64_t n, uint64_t *out,
				       unsigned long long *status, void *stream)
{
	if (n == 0 || n > NA_N)
		return -(int)hipErrorInvalidValue;
	void *base = nullptr;
	int rc = kmad_det_scratch_setup(&base);
	if (rc != 0)
		return rc;
	unsigned char *sp = (unsigned char *)base;
	kmad_det_bufs b;
	b.cls = (int8_t *)(sp + NA_CLS);
	b.rank = (uint32_t *)(sp + NA_RANK);
	b.table = (uint32_t *)(sp + NA_TABLE);
	b.evc = (uint32_t *)(sp + NA_EVC);
	b.evq = (uint32_t *)(sp + NA_EVQ);
	b.popped = (uint32_t *)(sp + NA_POPPED);
	b.cnt = (uint32_t *)(sp + NA_CNT);
	b.evcnt = (uint32_t *)(sp + NA_EVCNT);
	b.tot = (uint32_t *)(sp + NA_TOT);
	b.small = (uint32_t *)(sp + NA_SMALL);
	hipStream_t s = (hipStream_t)stream;
	/* the previous call's counters are still there: clear the small words
	 * and the caller's status with a kernel */
	hipLaunchKernelGGL(k_det_zero, dim3(1), dim3(64), 0, s, b.small, 64u, status, 2u);
	KMA_CHECK(hipGetLastError());
	rc = kmad_det_malloc(heap, sizes, n, &b, out, status, stream);
	if (rc != 0)
		return rc;
	hipLaunchKernelGGL(k_det_status, dim3(1), dim3(64), 0, s, b.small, status);
	KMA_CHECK(hipGetLastError());
	return 0;
}
