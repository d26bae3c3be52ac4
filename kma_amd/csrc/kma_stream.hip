/*
 * kma_stream.hip - gfx950 kernels that run a batched request stream
 * (include/kma_work.h, kma_stream_*) on a native heap:
 *   k_stream_plain      one launch per batch, per-lane kma_malloc / kma_free;
 *   k_stream_persistent one launch for the whole stream: work-groups claim
 *                       chunks of consecutive requests, serve the chunk's
 *                       mallocs slice by slice (256 requests) with
 *                       kma_wg_malloc, then its frees with kma_wg_free, at
 *                       the aggregation width the variant selects.
 *
 * Every malloc writes a pattern derived from its request index into the
 * block's first word and then publishes (epoch << 40 | heap offset) in the
 * result slot of its request.  A free reads its malloc's slot (waiting while
 * the slot still holds an older epoch), checks the pattern and frees the
 * block.  Frees only name requests of earlier batches, i.e. lower request
 * indices, and a chunk runs all its mallocs before any of its frees, so a
 * free only ever waits for a malloc pass that waits on nothing: the
 * persistent kernel always drains.  Pattern and result are written with
 * agent-scope atomic stores, and the pattern store completes before the
 * result store is issued, so a free running on another XCD sees both.
 */
#include <hip/hip_runtime.h>

#include "kma_device.h"
#include "kma_launch.h"
#include "kma_wg_device.h"
#include "kma_work.h"

#define KMA_CHECK(x)                                   \
	do {                                           \
		hipError_t e_ = (x);                   \
		if (e_ != hipSuccess)                  \
			return -(int)e_;               \
	} while (0)

namespace {

constexpr int SBLOCK = KMA_STREAM_BLOCK;
/* per-thread counters, summed per wavefront into ctl[KMA_CTL_STATS + k] */
enum { C_OPS, C_MOK, C_NULL, C_FOK, C_FNULL, C_BAD, C_TMO, C_LARGE, C_LNULL, C_N };
constexpr uint64_t OFF_MASK = (1ull << 40) - 1ull;
/* a free gives up waiting for its malloc's result after 2 s of wall clock */
constexpr uint64_t RESULT_WAIT_TICKS = 200000000ull;

/* Named fields, not an array: the compiler merges increments on different
 * paths into one with a computed index, which would put an array in scratch. */
/* Wavefront totals: every lane of the wavefront counts together, in uniform
 * control flow, with ballots, so the totals are scalars (kept out of the
 * vector registers the allocator's calls need). */
struct counters {
	uint32_t ops = 0, mok = 0, nul = 0, fok = 0, fnull = 0, bad = 0, tmo = 0, large = 0, lnull = 0;
};

__device__ __forceinline__ uint32_t wcount(bool b) { return (uint32_t)__popcll(__ballot(b)); }

__device__ __forceinline__ uint32_t pattern(uint64_t j)
{
	return (uint32_t)(j * 0x9E3779B1ull) ^ 0xA5A5A5A5u;
}

/* Per-wavefront phase clock (KMA_VAR_PHASES): wall-clock ticks spent in
 * malloc, in waiting for / checking free targets, and in free. */
struct phases {
	uint64_t t = 0, acc[3] = {0, 0, 0}, n = 0;
	bool on;
	__device__ explicit phases(bool enable) : on(enable) {}
	__device__ __forceinline__ void start()
	{
		if (on)
			t = wall_clock64();
	}
	__device__ __forceinline__ void mark(int k)
	{
		if (on) {
			uint64_t u = wall_clock64();
			acc[k] += u - t;
			t = u;
		}
	}
};

/* all 64 lanes of the wavefront are active when this runs */
__device__ __forceinline__ void flush(uint64_t *ctl, const counters &c, const phases &ph)
{
	const uint32_t v[C_N] = {c.ops, c.mok, c.nul, c.fok, c.fnull, c.bad, c.tmo, c.large, c.lnull};
#pragma unroll
	for (int k = 0; k < C_N; k++)
		if (__lane_id() == 0 && v[k])
			atomicAdd((unsigned long long *)(ctl + KMA_CTL_STATS + k), (unsigned long long)v[k]);
	if (ph.on && __lane_id() == 0) {
		for (int k = 0; k < 3; k++)
			atomicAdd((unsigned long long *)(ctl + KMA_CTL_STATS + C_N + k), (unsigned long long)ph.acc[k]);
		atomicAdd((unsigned long long *)(ctl + KMA_CTL_STATS + C_N + 3), (unsigned long long)ph.n);
	}
}

/* every lane calls; ism: the lane made a malloc of `size` that returned p */
__device__ __forceinline__ void count_malloc(const kma_dheap *H, bool ism, uint32_t size, const void *p, counters &c)
{
	const bool large = ism && size > H->max_small;
	c.mok += wcount(ism && p != nullptr);
	c.nul += wcount(ism && p == nullptr);
	c.large += wcount(large);
	c.lnull += wcount(large && !p);
}

/* Outcome of a free's wait for its malloc's result. */
enum { T_OK, T_BAD, T_NULL, T_TIMEOUT };

/* The block of request t's malloc: waits for the result slot to carry this
 * run's epoch and checks the pattern.  NULL if the malloc failed or the wait
 * gave up; *kind says which (counted by the caller with arithmetic only: an
 * increment per branch here was merged by the compiler into one through a
 * selected address, which put the counters in scratch). */
__device__ __forceinline__ void *take_target(const kma_dheap *H, uint64_t *res, uint32_t t, uint32_t ep, uint32_t *kind)
{
	using namespace kma_nat;
	uint64_t v = ld(res + t);
	if ((v >> 40) != ep) {
		uint64_t t0 = wall_clock64();
		uint32_t bo = 1;
		for (;;) {
			for (uint32_t i = 0; i < bo; i++)
				__builtin_amdgcn_s_sleep(2);
			bo = bo < 32u ? 2u * bo : 32u;
			v = fadd(res + t, (uint64_t)0);
			if ((v >> 40) == ep)
				break;
			if (wall_clock64() - t0 > RESULT_WAIT_TICKS) {
				*kind = T_TIMEOUT;
				return nullptr;
			}
		}
	}
	uint64_t off = v & OFF_MASK;
	if (!off) {
		*kind = T_NULL;
		return nullptr;
	}
	char *p = base(H) + off;
	*kind = ld((uint32_t *)p) != pattern(t) ? T_BAD : T_OK;
	return p;
}

/* every lane calls; isf: the lane waited for a free's target, with outcome kind */
__device__ __forceinline__ void count_free(bool isf, uint32_t kind, counters &c)
{
	c.fok += wcount(isf && (kind == T_OK || kind == T_BAD));
	c.bad += wcount(isf && kind == T_BAD);
	c.fnull += wcount(isf && kind == T_NULL);
	c.tmo += wcount(isf && kind == T_TIMEOUT);
}

__device__ __forceinline__ uint64_t result_word(const kma_dheap *H, const void *p, uint32_t ep)
{
	return ((uint64_t)ep << 40) | (p ? (uint64_t)((const char *)p - kma_nat::base(H)) : 0ull);
}

} // namespace

/* Start of a replay: next epoch (never 0, the value of a cleared slot),
 * chunk counter and counters to zero. */
__global__ void k_stream_begin(uint64_t *ctl)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	uint64_t e = (ctl[KMA_CTL_EPOCH] + 1u) & 0xffffffu;
	ctl[KMA_CTL_EPOCH] = e ? e : 1u;
	ctl[KMA_CTL_NEXT] = 0;
	for (int k = 0; k < 16; k++)
		ctl[KMA_CTL_STATS + k] = 0;
}

__global__ void __launch_bounds__(SBLOCK) k_stream_plain(kma_dheap *H, const uint32_t *req, uint64_t *res, uint64_t lo,
							 uint64_t hi, uint64_t *ctl, uint64_t *ts, uint32_t variant)
{
	using namespace kma_nat;
	const uint32_t ep = (uint32_t)ctl[KMA_CTL_EPOCH];
	const uint64_t j = lo + (uint64_t)blockIdx.x * SBLOCK + threadIdx.x;
	const bool log = (variant & KMA_VAR_LOG) != 0;
	counters c = {};
	phases ph((variant & KMA_VAR_PHASES) != 0);
	const bool in = j < hi;
	const uint32_t r = in ? req[j] : 0u;
	const bool isf = in && (r >> 31);
	const bool ism = in && !isf;
	void *p = nullptr;
	ph.start();
	if (ism) {
		p = kma_malloc(H, r);
		if (p)
			st((uint32_t *)p, pattern(j));
	}
	count_malloc(H, ism, r, p, c);
	order_prior(); /* pattern stores complete before results publish */
	if (ism) {
		st(res + j, result_word(H, p, ep));
		if (log)
			ts[j] = wall_clock64();
	}
	ph.mark(0);
	void *q = nullptr;
	uint32_t kind = T_OK;
	if (isf) {
		q = take_target(H, res, r & 0x7fffffffu, ep, &kind);
		if (log)
			ts[j] = wall_clock64();
	}
	count_free(isf, kind, c);
	ph.mark(1);
	if (isf)
		kma_free(H, q);
	ph.mark(2);
	ph.n = 1;
	c.ops = wcount(in);
	flush(ctl, c, ph);
}

template <uint32_t AGG>
__global__ void __launch_bounds__(SBLOCK) k_stream_persistent(kma_dheap *H, const uint32_t *req, uint64_t *res,
							      uint64_t nreq, uint32_t per_thread, uint64_t *ctl,
							      uint64_t *ts, uint32_t variant)
{
	using namespace kma_nat;
	__shared__ kma_wg_lds<SBLOCK> L;
	__shared__ uint64_t s_base;
	kma_wg_lds_init(L);
	const uint32_t ep = (uint32_t)ctl[KMA_CTL_EPOCH];
	const uint32_t tid = threadIdx.x;
	const bool log = (variant & KMA_VAR_LOG) != 0;
	const uint64_t chunk = (uint64_t)SBLOCK * per_thread;
	counters c = {};
	phases ph((variant & KMA_VAR_PHASES) != 0);
	for (;;) {
		if (tid == 0)
			s_base = fadd(ctl + KMA_CTL_NEXT, chunk);
		__syncthreads();
		const uint64_t b0 = s_base;
		__syncthreads();
		if (b0 >= nreq)
			break;
		/* every malloc of the chunk first, then its frees: a malloc never
		 * waits behind a free, so a free waits at most for other chunks'
		 * malloc passes and waits cannot chain */
		const uint32_t nsl = (uint32_t)min((uint64_t)per_thread, (nreq - b0 + SBLOCK - 1) / SBLOCK);
		ph.start();
		for (uint32_t m = 0; m < nsl; m++) {
			const uint64_t j = b0 + (uint64_t)m * SBLOCK + tid;
			const bool in = j < nreq;
			const uint32_t r = in ? req[j] : 0u;
			const uint32_t size = (in && !(r >> 31)) ? r : 0u;
			void *p = kma_wg_malloc<SBLOCK, AGG>(H, size, L);
			count_malloc(H, size != 0u, size, p, c);
			if (size && p)
				st((uint32_t *)p, pattern(j));
			order_prior();
			if (size) {
				st(res + j, result_word(H, p, ep));
				if (log)
					ts[j] = wall_clock64();
			}
			c.ops += wcount(in);
		}
		ph.mark(0);
		for (uint32_t m = 0; m < nsl; m++) {
			const uint64_t j = b0 + (uint64_t)m * SBLOCK + tid;
			const bool in = j < nreq;
			const uint32_t r = in ? req[j] : 0u;
			void *q = nullptr;
			uint32_t kind = T_OK;
			const bool isf = in && (r >> 31);
			if (isf) {
				q = take_target(H, res, r & 0x7fffffffu, ep, &kind);
				if (log)
					ts[j] = wall_clock64();
			}
			count_free(isf, kind, c);
			ph.mark(1);
			kma_wg_free<SBLOCK, AGG>(H, q, L);
			ph.mark(2);
		}
		ph.n += nsl;
	}
	flush(ctl, c, ph);
}

extern "C" int kmad_stream_begin(uint64_t *ctl, void *stream)
{
	hipLaunchKernelGGL(k_stream_begin, dim3(1), dim3(64), 0, (hipStream_t)stream, ctl);
	KMA_CHECK(hipGetLastError());
	return 0;
}

extern "C" int kmad_stream_plain(void *heap, const uint32_t *req, uint64_t *res, uint64_t lo, uint64_t n,
				 uint64_t *ctl, uint64_t *ts, uint32_t variant, void *stream)
{
	if (n == 0)
		return 0;
	uint64_t blocks = (n + SBLOCK - 1) / SBLOCK;
	hipLaunchKernelGGL(k_stream_plain, dim3((uint32_t)blocks), dim3(SBLOCK), 0, (hipStream_t)stream,
			   (kma_dheap *)heap, req, res, lo, lo + n, ctl, ts, variant);
	KMA_CHECK(hipGetLastError());
	return 0;
}

extern "C" int kmad_stream_persistent(void *heap, const uint32_t *req, uint64_t *res, uint64_t nreq, uint32_t nwg,
				      uint32_t per_thread, uint64_t *ctl, uint64_t *ts, uint32_t variant,
				      void *stream)
{
	if (nwg == 0 || per_thread == 0)
		return -1;
	hipStream_t s = (hipStream_t)stream;
	kma_dheap *H = (kma_dheap *)heap;
	switch ((variant & KMA_VAR_AGG_MASK) >> KMA_VAR_AGG_SHIFT) {
	case 1:
		hipLaunchKernelGGL(k_stream_persistent<8u>, dim3(nwg), dim3(SBLOCK), 0, s, H, req, res, nreq, per_thread,
				   ctl, ts, variant);
		break;
	case 2:
		hipLaunchKernelGGL(k_stream_persistent<16u>, dim3(nwg), dim3(SBLOCK), 0, s, H, req, res, nreq, per_thread,
				   ctl, ts, variant);
		break;
	case 3:
		hipLaunchKernelGGL(k_stream_persistent<32u>, dim3(nwg), dim3(SBLOCK), 0, s, H, req, res, nreq, per_thread,
				   ctl, ts, variant);
		break;
	default:
		hipLaunchKernelGGL(k_stream_persistent<KMA_AGG_WAVE>, dim3(nwg), dim3(SBLOCK), 0, s, H, req, res, nreq,
				   per_thread, ctl, ts, variant);
	}
	KMA_CHECK(hipGetLastError());
	return 0;
}
