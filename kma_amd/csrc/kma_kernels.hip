/*
 * kma_kernels.hip - gfx950 kernels behind libkma.so: heap initialisation,
 * serialized traces and the per-launch ("plain") workloads.
 */
#include <hip/hip_runtime.h>

#include "kma_device.h"
#include "kma_launch.h"
#include "kma_layout.h"
#include "kma_ref_device.h"
#include "kma_wg_device.h"
#include "kma_work.h"
#include "kma_workloads.h"

/* the host mirror of the device header must match it byte for byte */
static_assert(sizeof(kma_dheap) == sizeof(nat_hdr), "kma_dheap / nat_hdr size");
static_assert(offsetof(kma_dheap, off_psum) == offsetof(nat_hdr, off_psum), "off_psum");
static_assert(offsetof(kma_dheap, spare_slots) == offsetof(nat_hdr, spare_slots), "spare_slots");
static_assert(offsetof(kma_dheap, class_magic) == offsetof(nat_hdr, class_magic), "class_magic");
static_assert(KMA_STATS_BYTES == NAT_STATS_BYTES && KMA_PAGE_LEVELS == NAT_PAGE_LEVELS, "layout constants");

#define KMA_CHECK(x)                                   \
	do {                                           \
		hipError_t e_ = (x);                   \
		if (e_ != hipSuccess)                  \
			return -(int)e_;               \
	} while (0)

/* ------------------------------------------------------------------ */
/* reference layout: parallel heap init                                 */
/* ------------------------------------------------------------------ */
/*
 * The reference's clheap_init (kma.cl:27-52) enqueues superblocks 1..S-2
 * one by one behind the dummy sb[0] from a single work-item.  On a
 * zero-filled buffer every word that chain of enqueues writes has a closed
 * form, so each thread writes one superblock's link word directly:
 *   sb[0].next   = PTR(idx(sb1), 1)            (tag 1 from the queue init)
 *   sb[i].next   = PTR(idx(sb[i+1]), 0)        for 1 <= i <= S-3
 *   sb[S-2].next = 0
 *   head = PTR(idx(sb0), 0),  tail = PTR(idx(sb[S-2]), S-2)
 * where PTR(i, t) = (i & 0x1fffe) | ((t+1) & 0x7ff) << 22 and
 * idx(sb[k]) = (k+1) << 1 (clIndexedQueue.cl:44-65,89).
 */
__global__ void k_ref_init(char *heap, uint32_t log2, uint32_t S)
{
	using namespace kma_ref;
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t buckets = log2 - 1u;
	uint32_t hdr = 32u + 8u * buckets;
	char *sb0 = heap + hdr;
	if (i == 0) {
		uint64_t bytes = ((uint64_t)S << log2) + hdr;
		*reinterpret_cast<uint64_t *>(heap + 0) = bytes;
		*reinterpret_cast<uint64_t *>(heap + 8) = (uint64_t)(uintptr_t)sb0;
		*reinterpret_cast<uint64_t *>(heap + 16) = log2;
		*reinterpret_cast<uint32_t *>(heap + 24) = q_ptr(2u, 0u);
		*reinterpret_cast<uint32_t *>(heap + 28) = q_ptr(2u * (S - 1u), S - 2u);
		for (uint32_t c = 0; c < buckets; c++)
			*reinterpret_cast<uint64_t *>(heap + 32 + 8 * c) = 0;
	}
	if (i + 1 < S) { /* nodes sb[0..S-2] */
		uint32_t w;
		if (i + 2 == S)
			w = (i == 0) ? q_ptr(0u, 0u) : 0u;
		else if (i == 0)
			w = q_ptr(4u, 1u);
		else
			w = q_ptr((uint32_t)(2u * (i + 2u)), 0u);
		*reinterpret_cast<uint32_t *>(sb0 + (i << log2)) = w;
	}
}

extern "C" int kmad_ref_init(void *heap, uint32_t log2, uint32_t sblocks)
{
	uint32_t threads = 256;
	uint32_t blocks = (sblocks + threads - 1) / threads;
	hipLaunchKernelGGL(k_ref_init, dim3(blocks), dim3(threads), 0, 0, (char *)heap, log2, sblocks);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

__global__ void k_ref_arm(clheap *heap, uint64_t budget)
{
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		kma_ref::view v = kma_ref::make_view(heap);
		uint64_t dl = budget ? wall_clock64() + budget : 0;
		__hip_atomic_store(v.trl + kma_ref::T_DEADLINE, dl, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
	}
}

extern "C" int kmad_ref_arm(void *heap, uint64_t budget_ticks, void *stream)
{
	hipLaunchKernelGGL(k_ref_arm, dim3(1), dim3(64), 0, (hipStream_t)stream, (clheap *)heap, budget_ticks);
	KMA_CHECK(hipGetLastError());
	return 0;
}

/* ------------------------------------------------------------------ */
/* native layout: metadata init                                         */
/* ------------------------------------------------------------------ */
__global__ void k_nat_init(kma_dheap *H, uint64_t n)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		kma_nat::init_item(H, i);
}

extern "C" int kmad_nat_init(void *heap, uint64_t nsb, uint32_t npool, uint64_t nheads, uint64_t npbwords)
{
	(void)nheads;
	uint64_t n = nsb > npool ? nsb : npool;
	uint64_t nps = KMA_PAGE_LEVELS * ((npbwords + 63) / 64);
	if (nps > n)
		n = nps;
	if (npool % KMA_NXCD || npool / KMA_NXCD < 2)
		return -1;
	uint32_t threads = 256;
	uint64_t blocks = (n + threads - 1) / threads;
	hipLaunchKernelGGL(k_nat_init, dim3((uint32_t)blocks), dim3(threads), 0, 0, (kma_dheap *)heap, n);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

/* Detach every context head (no allocation may be in flight). */
__global__ void k_trim(kma_dheap *H, uint32_t nctx, uint32_t nclasses)
{
	using namespace kma_nat;
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nctx * nclasses)
		return;
	uint32_t ctx = i / nclasses, c = i % nclasses;
	uint64_t h = xchg(head(H, ctx, c), HEAD_EMPTY);
	if (h <= HEAD_BUSY)
		return;
	uint32_t sb = (uint32_t)h - 2u;
	uint64_t s = f_and(state(H, sb), ~ST_OWNED) & ~ST_OWNED;
	after_detach(H, sb, s, st_home(s), ctx);
}

/* Partial index (no allocation may be in flight): every listed superblock
 * that holds no live block is taken out of the index and released (into a
 * dirty shard of its home or a spare line, drained by k_clean_dirty).  One
 * thread per level-0 word; then k_pidx_rebuild, one thread per (XCD,
 * class), rebuilds the summary levels. */
__global__ void k_trim_pidx(kma_dheap *H, uint32_t nwords)
{
	using namespace kma_nat;
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nwords)
		return;
	const uint32_t w0 = pidx_w0(H), i0 = i % w0, c = (i / w0) % H->nclasses, x = i / (w0 * H->nclasses);
	trim_pidx_word(H, x, c, i0);
}

__global__ void k_pidx_rebuild(kma_dheap *H)
{
	using namespace kma_nat;
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < KMA_NXCD * H->nclasses)
		pidx_rebuild(H, i / H->nclasses, i % H->nclasses);
}

/* After a kernel boundary: drain dirty shards and spare lines into clean
 * shards (kma_nat::clean_all, one thread). */
__global__ void k_clean_dirty(kma_dheap *H)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	kma_nat::clean_all(H);
}

extern "C" int kmad_trim(void *heap, uint32_t nctx, uint32_t nclasses, uint32_t pidx_w0)
{
	uint32_t n = nctx * nclasses;
	hipLaunchKernelGGL(k_trim, dim3((n + 255) / 256), dim3(256), 0, 0, (kma_dheap *)heap, nctx, nclasses);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	uint32_t nw = KMA_NXCD * nclasses * pidx_w0;
	hipLaunchKernelGGL(k_trim_pidx, dim3((nw + 255) / 256), dim3(256), 0, 0, (kma_dheap *)heap, nw);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	hipLaunchKernelGGL(k_pidx_rebuild, dim3(1), dim3(256), 0, 0, (kma_dheap *)heap);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	hipLaunchKernelGGL(k_clean_dirty, dim3(1), dim3(64), 0, 0, (kma_dheap *)heap);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

/* ------------------------------------------------------------------ */
/* serialized trace: one work-item replays the op list                  */
/* ------------------------------------------------------------------ */
template <class HeapT>
__global__ void k_serial(HeapT *heap, const int64_t *ops, uint64_t n, uint64_t *out, const uint32_t *pat)
{
	if (blockIdx.x != 0 || threadIdx.x != 0)
		return;
	char *hb = reinterpret_cast<char *>(heap);
	for (uint64_t i = 0; i < n; i++) {
		int64_t op = ops[i];
		if (op >= 0) {
			char *p = (char *)kma_malloc(heap, (size_t)op);
			out[i] = p ? (uint64_t)(p - hb) : 0;
			if (p && pat)
				*reinterpret_cast<volatile uint32_t *>(p) = pat[i];
		} else {
			uint64_t off = out[(uint64_t)(-op - 1)];
			kma_free(heap, off ? (void *)(hb + off) : nullptr);
			out[i] = 0;
		}
	}
}

extern "C" int kmad_serial(int layout, void *heap, const int64_t *d_ops, uint64_t n, uint64_t *d_out,
			   const uint32_t *d_pat)
{
	if (layout == 0)
		hipLaunchKernelGGL(k_serial<clheap>, dim3(1), dim3(64), 0, 0, (clheap *)heap, d_ops, n, d_out, d_pat);
	else
		hipLaunchKernelGGL(k_serial<kma_dheap>, dim3(1), dim3(64), 0, 0, (kma_dheap *)heap, d_ops, n, d_out, d_pat);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

/* ------------------------------------------------------------------ */
/* plain per-launch workloads                                           */
/* ------------------------------------------------------------------ */
template <class HeapT, bool GENERIC>
__global__ void __launch_bounds__(256) k_rounds(HeapT *heap, kma_wl_args a)
{
	wl_rounds<HeapT, GENERIC>(heap, a);
}

extern "C" int kmad_rounds(int layout, void *heap, uint32_t round0, uint32_t nrounds, int mode, uint32_t fixed,
			   uint64_t seed, uint64_t nwi, uint32_t block, uint32_t variant, unsigned long long *d_stats,
			   uint32_t *d_lat, void *stream)
{
	if (block == 0 || block > 256 || (block & 63))
		return -1;
	uint64_t blocks = (nwi + block - 1) / block;
	hipStream_t s = (hipStream_t)stream;
	kma_wl_args a = {round0, nrounds, mode, fixed, seed, nwi, variant, d_stats, d_lat};
	const bool generic = variant != 0 || d_lat != nullptr;
	if (layout == 0 && generic)
		hipLaunchKernelGGL((k_rounds<clheap, true>), dim3((uint32_t)blocks), dim3(block), 0, s, (clheap *)heap, a);
	else if (layout == 0)
		hipLaunchKernelGGL((k_rounds<clheap, false>), dim3((uint32_t)blocks), dim3(block), 0, s, (clheap *)heap, a);
	else if (generic)
		hipLaunchKernelGGL((k_rounds<kma_dheap, true>), dim3((uint32_t)blocks), dim3(block), 0, s,
				   (kma_dheap *)heap, a);
	else
		hipLaunchKernelGGL((k_rounds<kma_dheap, false>), dim3((uint32_t)blocks), dim3(block), 0, s,
				   (kma_dheap *)heap, a);
	KMA_CHECK(hipGetLastError());
	return 0;
}

/* ------------------------------------------------------------------ */
/* roofline microbenchmarks                                             */
/* ------------------------------------------------------------------ */
/*
 * Agent-scope atomicOr throughput, the operation the allocator's bitmaps
 * use.  pattern 0: every lane hits a pseudo-random distinct 4-B word over
 * the whole buffer (distinct-address peak); 1: lanes of a wavefront hit
 * consecutive words (coalesced); 2: all lanes of a wavefront hit one word
 * (contended).  Each lane issues `iters` atomics.  Pattern 3 is the L2
 * request-rate reference for the full-occupancy workload's hold phase:
 * plain 4-B stores, each lane writing `iters` consecutive words of its own
 * 2 KiB slot at a pseudo-random position in the buffer.
 */
__global__ void __launch_bounds__(256) k_atomic_peak(uint32_t *buf, uint64_t nwords, uint32_t iters, int pattern)
{
	uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint64_t nth = (uint64_t)gridDim.x * blockDim.x;
	if (pattern == 3) {
		uint64_t slot = (kma_mix(7, gid, 0, 0) >> 8) % (nwords / 512u);
		uint32_t *p = buf + slot * 512u;
		for (uint32_t i = 0; i < iters; i++)
			wl_store(p + i % 511u, i);
		return;
	}
	uint32_t acc = 0;
	for (uint32_t i = 0; i < iters; i++) {
		uint64_t w;
		if (pattern == 0)
			w = kma_mix(7, gid, i, 0) % nwords;
		else if (pattern == 1)
			w = (gid + (uint64_t)i * nth) % nwords;
		else
			w = ((gid >> 6) * 64 + (uint64_t)i * 4099) % nwords;
		acc ^= __hip_atomic_fetch_or(buf + w, 1u << (i & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	}
	if (acc == 0x12345678u)
		buf[0] = acc; /* keep the results live */
}

/* Streaming copy of nvec 16-byte vectors (grid-stride): the reference
 * kernel for HBM bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE
 * counters (it reads and writes exactly nvec * 16 bytes). */
__global__ void __launch_bounds__(256) k_copy(const uint4 *src, uint4 *dst, uint64_t nvec)
{
	uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (; i < nvec; i += stride)
		dst[i] = src[i];
}

extern "C" int kmad_copy(const void *src, void *dst, uint64_t bytes, void *stream)
{
	uint64_t nvec = bytes / 16;
	hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, (hipStream_t)stream, (const uint4 *)src, (uint4 *)dst,
			   nvec);
	KMA_CHECK(hipGetLastError());
	return 0;
}

extern "C" int kmad_atomic_peak(void *buf, uint64_t nwords, uint64_t nthreads, uint32_t iters, int pattern,
				void *stream)
{
	uint32_t block = 256;
	uint64_t blocks = (nthreads + block - 1) / block;
	hipLaunchKernelGGL(k_atomic_peak, dim3((uint32_t)blocks), dim3(block), 0, (hipStream_t)stream,
			   (uint32_t *)buf, nwords, iters, pattern);
	KMA_CHECK(hipGetLastError());
	return 0;
}

/* ------------------------------------------------------------------ */
/* batched persistent form                                              */
/* ------------------------------------------------------------------ */
/*
 * One launch runs every round; the work-groups allocate and free their
 * requests collectively (kma_wg_device.h) with the aggregation width AGG.
 * Work-items past nwi take part with empty requests so control flow stays
 * uniform.
 */
template <bool GENERIC, uint32_t AGG>
__global__ void __launch_bounds__(KMA_BATCH_BLOCK) k_batched(kma_dheap *H, kma_wl_args a)
{
	__shared__ kma_wg_lds<KMA_BATCH_BLOCK> L;
	kma_wg_lds_init(L);
	wl_batched<GENERIC, AGG>(H, L, a);
}

template <bool GENERIC>
static void launch_batched(uint32_t agg, uint64_t blocks, hipStream_t s, kma_dheap *H, const kma_wl_args &a)
{
	switch (agg) {
	case 1:
		hipLaunchKernelGGL((k_batched<GENERIC, 8u>), dim3((uint32_t)blocks), dim3(KMA_BATCH_BLOCK), 0, s, H, a);
		break;
	case 2:
		hipLaunchKernelGGL((k_batched<GENERIC, 16u>), dim3((uint32_t)blocks), dim3(KMA_BATCH_BLOCK), 0, s, H, a);
		break;
	case 3:
		hipLaunchKernelGGL((k_batched<GENERIC, 32u>), dim3((uint32_t)blocks), dim3(KMA_BATCH_BLOCK), 0, s, H, a);
		break;
	default:
		hipLaunchKernelGGL((k_batched<GENERIC, KMA_AGG_WAVE>), dim3((uint32_t)blocks), dim3(KMA_BATCH_BLOCK), 0, s,
				   H, a);
	}
}

extern "C" int kmad_batched(void *heap, uint32_t round0, uint32_t nrounds, int mode, uint32_t fixed, uint64_t seed,
			    uint64_t nwi, uint32_t variant, unsigned long long *d_stats, uint32_t *d_lat, void *stream)
{
	uint64_t blocks = (nwi + KMA_BATCH_BLOCK - 1) / KMA_BATCH_BLOCK;
	const uint32_t agg = (variant & KMA_VAR_AGG_MASK) >> KMA_VAR_AGG_SHIFT;
	kma_wl_args a = {round0, nrounds, mode, fixed, seed, nwi, variant & ~KMA_VAR_AGG_MASK, d_stats, d_lat};
	if (a.variant == 0 && !d_lat)
		launch_batched<false>(agg, blocks, (hipStream_t)stream, (kma_dheap *)heap, a);
	else
		launch_batched<true>(agg, blocks, (hipStream_t)stream, (kma_dheap *)heap, a);
	KMA_CHECK(hipGetLastError());
	return 0;
}

/* ------------------------------------------------------------------ */
/* bitmap claim check: four cases per wavefront, one per 16-lane group  */
/* (the claim of the work-group collective malloc); case i uses          */
/* superblock i of a native heap preloaded with a bitmap                 */
/* ------------------------------------------------------------------ */
__global__ void __launch_bounds__(64) k_check_wave_claim(kma_dheap *H, uint32_t ncases, const uint32_t *cls,
							 const uint32_t *base, const uint32_t *got,
							 const uint32_t *bm_in, uint64_t *res_out, uint32_t *bm_out,
							 uint32_t *done_out)
{
	__shared__ uint16_t order[4][64];
	__shared__ uint64_t res[4][64];
	const uint32_t l = threadIdx.x, g = l / 16u;
	for (uint32_t q = 0; q < 4; q++) {
		uint32_t cs = blockIdx.x * 4u + q;
		if (cs < ncases)
			for (uint32_t w = l; w < H->bm_words; w += 64)
				kma_nat::bitmap(H, cs)[w] = bm_in[cs * H->bm_words + w];
		order[q][l] = (uint16_t)l;
		res[q][l] = 0;
	}
	__syncthreads();
	const uint32_t cs = blockIdx.x * 4u + g;
	const bool want = cs < ncases;
	uint32_t done = kma_nat::group_claim<16u>(H, want, cs, want ? cls[cs] : 0u, want ? base[cs] : 0u,
					     want ? got[cs] : 0u, order[g], res[g]);
	__syncthreads();
	for (uint32_t q = 0; q < 4; q++) {
		uint32_t c2 = blockIdx.x * 4u + q;
		if (c2 >= ncases)
			continue;
		res_out[c2 * 64 + l] = res[q][l];
		for (uint32_t w = l; w < H->bm_words; w += 64)
			bm_out[c2 * H->bm_words + w] = kma_nat::ld(kma_nat::bitmap(H, c2) + w);
	}
	if (want && kma_nat::grp_lane<16u>() == 0)
		done_out[cs] = done;
}

/* One wavefront frees offs[lane] (0: no request) in a single kma_free call. */
__global__ void __launch_bounds__(64) k_wave_free(kma_dheap *H, const uint64_t *offs, uint32_t n)
{
	const uint32_t l = threadIdx.x;
	const uint64_t o = l < n ? offs[l] : 0u;
	kma_free(H, o ? (void *)((char *)H + o) : nullptr);
}

extern "C" int kmad_wave_free(void *heap, const uint64_t *offs, uint32_t n)
{
	hipLaunchKernelGGL(k_wave_free, dim3(1), dim3(64), 0, 0, (kma_dheap *)heap, offs, n);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}

extern "C" int kmad_check_wave_claim(void *heap, uint32_t ncases, const uint32_t *cls, const uint32_t *base,
				     const uint32_t *got, const uint32_t *bm_in, uint64_t *res_out, uint32_t *bm_out,
				     uint32_t *done_out)
{
	hipLaunchKernelGGL(k_check_wave_claim, dim3((ncases + 3u) / 4u), dim3(64), 0, 0, (kma_dheap *)heap, ncases, cls,
			   base, got, bm_in, res_out, bm_out, done_out);
	KMA_CHECK(hipGetLastError());
	KMA_CHECK(hipDeviceSynchronize());
	return 0;
}
