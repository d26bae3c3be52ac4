/*
 * kma_workloads.h - device bodies of the per-work-item workloads (plain and
 * fused forms, per-lane API) and of the batched form (work-group collective
 * API).  Kept apart from the kernels so the same code also builds for the
 * CPU simulation used by the tests.
 */
#ifndef KMA_WORKLOADS_H
#define KMA_WORKLOADS_H

#include "kma_device.h"
#include "kma_ref_device.h"
#include "kma_wg_device.h"
#include "kma_work.h"

/* ------------------------------------------------------------------ */
/* deterministic trace values (same function as the host side)         */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ uint64_t kma_mix(uint64_t seed, uint64_t gid, uint64_t round, uint64_t k)
{
	uint64_t z = seed * 0x9E3779B97F4A7C15ull + gid * 0xBF58476D1CE4E5B9ull +
		     round * 0x94D049BB133111EBull + k * 0xD6E8FEB86659FD93ull;
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

/* Workload payload accesses are ordinary cached stores and loads.  The
 * empty asm with a memory clobber after each store keeps the compiler from
 * merging or dropping any of them (every hold iteration's write is issued),
 * and before the read-back it forces a real load.  (A volatile pointer would
 * make each access a system-scope write-through with a wait after it.) */
__device__ __forceinline__ void wl_store(uint32_t *p, uint32_t v)
{
	*p = v;
	__asm__ __volatile__("" ::: "memory");
}
__device__ __forceinline__ uint32_t wl_load(const uint32_t *p)
{
	__asm__ __volatile__("" ::: "memory");
	return *p;
}
__device__ __forceinline__ void wl_store64(uint64_t *p, uint64_t v)
{
	*p = v;
	__asm__ __volatile__("" ::: "memory");
}

__device__ __forceinline__ void wave_add_stats(unsigned long long *stats, uint64_t a, uint64_t b, uint64_t c,
					       uint64_t d)
{
	/* one atomic per counter per wavefront */
	uint64_t v[4] = {a, b, c, d};
	for (int k = 0; k < 4; k++) {
		uint64_t x = v[k];
		for (int off = 32; off > 0; off >>= 1)
			x += __shfl_xor(x, off);
		if (__lane_id() == 0 && x)
			atomicAdd(stats + k, (unsigned long long)x);
	}
}

/*
 * One kernel body for every per-work-item workload:
 *  KMA_WL_FIXED    : malloc(fixed), write {gid, round}, read back, free
 *                    (plain small workload)
 *  KMA_WL_LOGUNI   : malloc(2^e), e uniform in 3..11, hold 0..64 writes,
 *                    read back, free (full-occupancy workload)
 *  KMA_WL_REF_TEST : kma_test_malloc (kma.cl:424-461): malloc(8), write pid,
 *                    free; a NULL ends the work-item
 *  KMA_WL_LOWVAR   : kma_test_malloc_lowvar (kma.cl:463-492)
 *  KMA_WL_HIGHVAR  : kma_test_malloc_highvar (kma.cl:494-526)
 * stats: [0] heap ops issued, [1] NULL returns, [2] read-back mismatches,
 *        [3] mallocs issued.
 */
/* Launch arguments shared by the workload kernels. */
struct kma_wl_args {
	uint32_t round0, nrounds;
	int mode;
	uint32_t fixed;
	uint64_t seed;
	uint64_t nwi;
	uint32_t variant;          /* KMA_VAR_* bits */
	unsigned long long *stats; /* [0] ops [1] nulls [2] bad [3] mallocs
				      [4..6] malloc/hold/free ticks summed over
				      wavefronts and rounds, [7] samples */
	uint32_t *lat;             /* optional: malloc ticks per (round, work-item) */
};

/* Per-wavefront phase clock (lane 0's view; the wavefront runs in step). */
struct kma_phase {
	uint64_t t = 0, acc[3] = {0, 0, 0}, n = 0;
	bool on;
	__device__ kma_phase(bool enable) : on(enable) {}
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
			n += (k == 2);
		}
	}
	__device__ __forceinline__ void flush(unsigned long long *stats)
	{
		if (on && __lane_id() == 0) {
			for (int k = 0; k < 3; k++)
				atomicAdd(stats + 4 + k, (unsigned long long)acc[k]);
			atomicAdd(stats + 7, (unsigned long long)n);
		}
	}
};

/* GENERIC = false: variant 0 and no latency probes compiled in (see
 * wl_batched). */
template <class HeapT, bool GENERIC>
__device__ inline void wl_rounds(HeapT *heap, const kma_wl_args &a)
{
	uint32_t *const lat = GENERIC ? a.lat : nullptr;
	const uint32_t round0 = a.round0, nrounds = a.nrounds, fixed = a.fixed;
	const int mode = a.mode;
	const uint64_t seed = a.seed, nwi = a.nwi;
	unsigned long long *stats = a.stats;
	uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t ops = 0, nulls = 0, bad = 0, mallocs = 0; /* per work-item: a few per round */
	bool alive = gid < nwi;
	kma_phase ph(GENERIC && (a.variant & KMA_VAR_PHASES) != 0);
	for (uint32_t r = round0; r < round0 + nrounds; r++) {
		if (!alive)
			break;
		if (mode == KMA_WL_FIXED || mode == KMA_WL_LOGUNI) {
			uint32_t size, hold = 0;
			if (mode == KMA_WL_FIXED) {
				size = fixed;
			} else {
				size = 1u << (3 + kma_mix(seed, gid, r, 0) % 9);
				hold = (uint32_t)(kma_mix(seed, gid, r, 1) % 65);
			}
			ph.start();
			uint64_t t0 = lat ? wall_clock64() : 0;
			uint32_t *p = (uint32_t *)kma_malloc(heap, size);
			if (lat)
				lat[(uint64_t)(r - round0) * nwi + gid] = (uint32_t)(wall_clock64() - t0);
			ph.mark(0);
			mallocs++;
			ops++;
			if (!p) {
				nulls++;
				continue;
			}
			uint32_t pat = (uint32_t)(gid << 8) ^ r;
			if (mode == KMA_WL_FIXED) {
				wl_store(p, (uint32_t)gid);
				wl_store(p + 1, r);
				if (wl_load(p) != (uint32_t)gid || wl_load(p + 1) != r)
					bad++;
			} else {
				uint32_t words = size >> 2;
				wl_store(p, pat);
				for (uint32_t k = 0; k < hold; k++)
					wl_store(p + 1 + k % (words - 1), pat + k + 1);
				if (wl_load(p) != pat)
					bad++;
			}
			ph.mark(1);
			kma_free(heap, (void *)p);
			ph.mark(2);
			ops++;
		} else if (mode == KMA_WL_REF_TEST) {
			uint64_t *p = (uint64_t *)kma_malloc(heap, 8);
			mallocs++;
			ops++;
			if (!p) {
				nulls++;
				alive = false;
				continue;
			}
			wl_store64(p, gid);
			kma_free(heap, (void *)p);
			ops++;
		} else {
			uint32_t amount;
			if (mode == KMA_WL_LOWVAR)
				amount = (gid & 1) ? 8 : 16;
			else
				amount = 4u << ((gid + r) % 5);
			uint32_t *p = (uint32_t *)kma_malloc(heap, amount);
			mallocs++;
			ops++;
			if (!p) {
				nulls++;
				alive = false;
				continue;
			}
			wl_store(p, mode == KMA_WL_LOWVAR ? (uint32_t)gid : 4919u);
			if (gid == 0) {
				uint32_t *q = (uint32_t *)kma_malloc(heap, 4000);
				mallocs++;
				ops++;
				if (!q) {
					nulls++;
					alive = false;
					continue;
				}
				wl_store(q + 8, 4919u);
				kma_free(heap, (void *)q);
				ops++;
			}
			kma_free(heap, (void *)p);
			ops++;
		}
	}
	wave_add_stats(stats, ops, nulls, bad, mallocs);
	ph.flush(stats);
}

#define KMA_BATCH_BLOCK 256

/* GENERIC = false compiles the plain collective path only (variant 0, no
 * latency probes): the per-lane malloc/free and the probes would otherwise
 * set the register budget, and with it the occupancy, of the whole kernel. */
template <bool GENERIC, uint32_t AGG>
__device__ inline void wl_batched(kma_dheap *H, kma_wg_lds<KMA_BATCH_BLOCK> &L, const kma_wl_args &a)
{
	const uint32_t round0 = a.round0, nrounds = a.nrounds, fixed = a.fixed;
	const uint32_t variant = GENERIC ? a.variant : 0u;
	uint32_t *const lat = GENERIC ? a.lat : nullptr;
	const int mode = a.mode;
	const uint64_t seed = a.seed, nwi = a.nwi;
	unsigned long long *stats = a.stats;
	kma_phase ph((variant & KMA_VAR_PHASES) != 0);
	uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	bool active = gid < nwi;
	uint32_t ops = 0, nulls = 0, bad = 0, mallocs = 0; /* per work-item: two per round */
	for (uint32_t r = round0; r < round0 + nrounds; r++) {
		uint32_t size = 0, hold = 0;
		if (active) {
			if (mode == KMA_WL_FIXED) {
				size = fixed;
			} else {
				size = 1u << (3 + kma_mix(seed, gid, r, 0) % 9);
				hold = (uint32_t)(kma_mix(seed, gid, r, 1) % 65);
			}
		}
		uint32_t *p;
		ph.start();
		uint64_t t0 = lat ? wall_clock64() : 0;
		if (variant & 1u)
			p = (uint32_t *)kma_malloc(H, size);
		else
			p = (uint32_t *)kma_wg_malloc<KMA_BATCH_BLOCK, AGG>(H, size, L, (variant >> 2) & 1u);
		if (lat && active)
			lat[(uint64_t)(r - round0) * nwi + gid] = (uint32_t)(wall_clock64() - t0);
		ph.mark(0);
		if (active) {
			mallocs++;
			ops++;
			if (!p) {
				nulls++;
			} else {
				uint32_t pat = (uint32_t)(gid << 8) ^ r;
				if (mode == KMA_WL_FIXED) {
					wl_store(p, (uint32_t)gid);
					wl_store(p + 1, r);
					if (wl_load(p) != (uint32_t)gid || wl_load(p + 1) != r)
						bad++;
				} else {
					uint32_t words = size >> 2;
					wl_store(p, pat);
					for (uint32_t k = 0; k < hold; k++)
						wl_store(p + 1 + k % (words - 1), pat + k + 1);
					if (wl_load(p) != pat)
						bad++;
				}
				ops++;
			}
		}
		ph.mark(1);
		if (variant & 2u)
			kma_free(H, (void *)p);
		else
			kma_wg_free<KMA_BATCH_BLOCK, AGG>(H, (void *)p, L);
		ph.mark(2);
	}
	wave_add_stats(stats, ops, nulls, bad, mallocs);
	ph.flush(stats);
}


#endif /* KMA_WORKLOADS_H */
