/*
 * kma_serve.hip - the device-resident serving loop (kma_serve_run in
 * include/kma_work.h): one cooperative persistent kernel that admits jobs as
 * they arrive, grows every live job by one block per step and frees a job's
 * blocks when it completes, on a native heap.
 *
 * One step = three phases separated by grid barriers:
 *   A (one lane of work-group 0): retire the jobs completed in the previous
 *     step (slot back to the free-slot stack), then admit arrived jobs in
 *     arrival order while the heap's pooled superblocks leave use below the
 *     watermark and a slot is free; decide whether the run is over;
 *   B (all work-groups): work-group g serves job slots [256 k, 256 k + 256)
 *     for k = g, g + G, ...: ONE collective kma_wg_malloc of block_bytes for
 *     every live slot, the block recorded in the slot's block table and
 *     stamped with pattern(job, step); a job that reaches its length appends
 *     (slot, job, first, count) to the done list, whose 64-bit counter keeps
 *     the entry index in bits 40.. and the running block total below, so
 *     entries are ordered by `first`;
 *   C (all work-groups): the blocks of all completed jobs, numbered
 *     0 .. total-1, are spread over the whole grid: each thread finds its
 *     entry by binary search, checks the block's pattern, adds it to the
 *     job's result and frees the block with kma_wg_free.
 * A NULL from malloc stalls that job for the step (retried next step).  After
 * a step with stalls, phase A preempts the most recently admitted live jobs
 * until their blocks cover the shortfall (their blocks are freed in phase C,
 * without adding to their results) and puts them on a requeue ring that is
 * admitted before new arrivals.  A step with stalls also pauses admission
 * until use falls SV_HYST_PCT points below the watermark (hysteresis), so
 * the freed room is not refilled at once with jobs that the next steps'
 * growth would preempt again.  Both keep the loop from filling the heap
 * with jobs that can then never grow (admission by a use watermark alone
 * cannot know how long jobs grow).
 * Barriers release (write back L2) and acquire (invalidate) at agent scope,
 * so plain stores of one phase are visible to every XCD in the next.  Each
 * barrier wait has a 5 s watchdog that aborts the whole loop.
 */
#include <hip/hip_runtime.h>

#include "kma_device.h"
#include "kma_launch.h"
#include "kma_wg_device.h"

#define KMA_CHECK(x)                                   \
	do {                                           \
		hipError_t e_ = (x);                   \
		if (e_ != hipSuccess)                  \
			return -(int)e_;               \
	} while (0)

#define SV_BLOCK 256
/* aggregation width of the loop's collective calls: work-group staging
 * with 16-lane claim groups (every work-group asks for 256 blocks of one
 * class per step, which whole-superblock grabs serve) */
#define SV_AGG 16u

namespace {

/* ctl indices, each on its own 128-B line */
enum : uint32_t {
	C_BAR_CNT = 0,
	C_BAR_GEN = 16,
	C_NEXT = 32,  /* next job to admit */
	C_NFREE = 48, /* free-slot stack top */
	C_STOP = 64,
	C_DONE = 80, /* done list: entries << 40 | blocks */
	C_T0 = 96,
	C_ABORT = 112,
	C_LIVE = 128,
	C_LTOP = 192,      /* phase A only: live-stack top, requeue head/tail, sequence */
	C_RQ_HEAD = 193,
	C_RQ_TAIL = 194,
	C_SEQ = 195,
	C_HOLD = 196,      /* admission paused since a step with stalls */
	C_POOLED = 224,      /* free superblocks, summed over the work-groups' slices */
	C_STEP_STALLS = 240, /* stalls of the current step (all work-groups) */
};
constexpr uint64_t PFLAG = 1ull << 63; /* done-list slot field: preempted, not completed */
/* counters at KMA_SV_STATS + k */
enum : uint32_t {
	S_STEPS,
	S_MALLOCS,
	S_STALLS,
	S_FREES,
	S_BAD,
	S_ADMITTED,
	S_DONE,
	S_PEAK_USED,
	S_RUN_TICKS,
	S_DRAINED,
	S_DRAIN_FREES,
	S_PREEMPT,
	S_N
};
constexpr uint64_t MASK40 = (1ull << 40) - 1ull;
constexpr uint32_t SV_HYST_PCT = 10; /* admission hysteresis, points below the watermark */
constexpr uint64_t BARRIER_TICKS = 500000000ull; /* 5 s */

__device__ __forceinline__ uint64_t pattern(uint64_t job, uint64_t k)
{
	uint64_t z = job * 0x9E3779B97F4A7C15ull + k * 0xBF58476D1CE4E5B9ull + 0x94D049BB133111EBull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t pattern_id(const kma_serve_args &a, uint64_t job)
{
	return a.job_id ? a.job_id[job] : job;
}

/* Grid barrier for a co-resident grid; false if the loop was aborted. */
__device__ __forceinline__ bool grid_sync(uint64_t *ctl, uint32_t nwg)
{
	using namespace kma_nat;
	__shared__ int s_ok;
	order_prior(); /* this thread's stores complete before the write-back */
	__syncthreads();
	if (threadIdx.x == 0) {
		int ok = 1;
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
		uint64_t gen = ld(ctl + C_BAR_GEN);
		uint64_t arrived = fadd(ctl + C_BAR_CNT, (uint64_t)1) + 1u;
		if (arrived == nwg) {
			st(ctl + C_BAR_CNT, (uint64_t)0);
			order_prior();
			fadd(ctl + C_BAR_GEN, (uint64_t)1);
		} else {
			uint64_t t0 = wall_clock64();
			while (ld(ctl + C_BAR_GEN) == gen) {
				__builtin_amdgcn_s_sleep(2);
				if (ld(ctl + C_ABORT)) {
					ok = 0;
					break;
				}
				if (wall_clock64() - t0 > BARRIER_TICKS) {
					st(ctl + C_ABORT, (uint64_t)1);
					ok = 0;
					break;
				}
			}
		}
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
		s_ok = ok && ld(ctl + C_ABORT) == 0;
	}
	__syncthreads();
	return s_ok != 0;
}

__device__ __forceinline__ bool lentry_live(const kma_serve_args &a, uint64_t e)
{
	uint32_t slot = (uint32_t)e;
	return a.slot_job[slot] >= 0 && a.slot_seq[slot] == (uint32_t)(e >> 32);
}

/* Phase A, one lane. */
__device__ __forceinline__ void admit(const kma_serve_args &a, uint64_t t0, uint64_t pooled)
{
	const kma_dheap *H = (const kma_dheap *)a.heap;
	uint64_t *ctl = a.ctl;
	uint64_t *S = ctl + KMA_SV_STATS;
	uint64_t now = wall_clock64() - t0;
	uint64_t nfree = ctl[C_NFREE], live = ctl[C_LIVE], next = ctl[C_NEXT];
	uint64_t ltop = ctl[C_LTOP], rqh = ctl[C_RQ_HEAD], rqt = ctl[C_RQ_TAIL], seq = ctl[C_SEQ];
	/* retire the previous step's entries: completions and preemptions */
	uint64_t dc = ctl[C_DONE];
	uint32_t ne = (uint32_t)(dc >> 40);
	for (uint32_t e = 0; e < ne; e++) {
		uint64_t sf = a.done[4u * e], job = a.done[4u * e + 1u];
		uint32_t slot = (uint32_t)(sf & ~PFLAG);
		if (sf & PFLAG) {
			a.requeue[rqt++ % a.rq_cap] = job;
		} else {
			a.job_done[job] = now + 1u;
			a.slot_job[slot] = -1;
			S[S_DONE]++;
			live--;
		}
		a.free_slots[nfree++] = slot;
	}
	dc = 0;
	uint64_t stalls = ctl[C_STEP_STALLS];
	ctl[C_STEP_STALLS] = 0;
	uint64_t used = H->nsb > pooled ? H->nsb - pooled : 0u;
	if (used > S[S_PEAK_USED])
		S[S_PEAK_USED] = used;
	bool stop = now >= a.duration_ticks || (next >= a.njobs && live == 0 && rqh == rqt);
	/* hysteresis: a step with stalls pauses admission until use falls
	 * SV_HYST_PCT points below the watermark, so the heap is not refilled
	 * with jobs that the next steps' growth would preempt again */
	bool hold = ctl[C_HOLD] != 0;
	if (stalls)
		hold = true;
	else if (hold && used * 100u < (uint64_t)(a.watermark_pct > SV_HYST_PCT ? a.watermark_pct - SV_HYST_PCT : 0u) * H->nsb)
		hold = false;
	ctl[C_HOLD] = hold ? 1u : 0u;
	if (!stop && stalls) {
		/* preempt the youngest live jobs until their blocks cover the
		 * shortfall; their blocks go to this step's free list */
		uint64_t need = stalls;
		while (need && ltop && rqt - rqh < a.rq_cap) {
			uint64_t e = a.lstack[--ltop];
			if (!lentry_live(a, e))
				continue;
			uint32_t slot = (uint32_t)e;
			uint64_t blocks = a.slot_steps[slot];
			uint64_t k = dc >> 40;
			a.done[4u * k] = slot | PFLAG;
			a.done[4u * k + 1u] = (uint64_t)a.slot_job[slot];
			a.done[4u * k + 2u] = dc & MASK40;
			a.done[4u * k + 3u] = blocks;
			dc += ((uint64_t)1 << 40) + blocks;
			a.slot_job[slot] = -1;
			live--;
			S[S_PREEMPT]++;
			need -= need < blocks ? need : blocks;
		}
	}
	ctl[C_DONE] = dc;
	if (!stop && !hold) {
		uint32_t k = 0;
		while (nfree > 0 && k < a.max_admit && used * 100u < (uint64_t)a.watermark_pct * H->nsb) {
			uint64_t job;
			if (rqh != rqt) {
				job = a.requeue[rqh++ % a.rq_cap];
			} else if (next < a.njobs && a.arrival[next] <= now) {
				job = next++;
				a.job_admit[job] = now + 1u;
			} else {
				break;
			}
			uint32_t slot = a.free_slots[--nfree];
			a.slot_job[slot] = (int64_t)job;
			a.slot_steps[slot] = 0;
			a.slot_seq[slot] = (uint32_t)++seq;
			if (ltop == a.lstack_cap) {
				/* drop dead entries; at most max_live stay */
				uint64_t w = 0;
				for (uint64_t r = 0; r < ltop; r++)
					if (lentry_live(a, a.lstack[r]))
						a.lstack[w++] = a.lstack[r];
				ltop = w;
			}
			a.lstack[ltop++] = (seq << 32) | slot;
			live++;
			k++;
		}
		S[S_ADMITTED] += k;
	}
	if (!stop)
		S[S_STEPS]++;
	else
		S[S_RUN_TICKS] = now;
	ctl[C_NFREE] = nfree;
	ctl[C_LIVE] = live;
	ctl[C_NEXT] = next;
	ctl[C_LTOP] = ltop;
	ctl[C_RQ_HEAD] = rqh;
	ctl[C_RQ_TAIL] = rqt;
	ctl[C_SEQ] = seq;
	ctl[C_STOP] = stop ? 1u : 0u;
}

} // namespace

__global__ void __launch_bounds__(SV_BLOCK) k_serve(kma_serve_args a)
{
	using namespace kma_nat;
	kma_dheap *H = (kma_dheap *)a.heap;
	__shared__ kma_wg_lds<SV_BLOCK> L;
	__shared__ uint32_t s_max;
	const uint32_t tid = threadIdx.x, nwg = gridDim.x;
	uint64_t *ctl = a.ctl;
	const uint32_t nslices = a.max_live / SV_BLOCK;
	uint64_t c_mallocs = 0, c_stalls = 0, c_frees = 0, c_bad = 0;
	if (blockIdx.x == 0 && tid == 0)
		ctl[C_T0] = wall_clock64();
	/* every work-group counts its slice of the heap's free superblocks
	 * (here, then after each step's frees); phase A reads the sum */
	if (tid < 64) {
		uint64_t part = kma_free_superblocks_part(H, blockIdx.x, nwg);
		if (tid == 0)
			fadd(ctl + C_POOLED, part);
	}
	if (!grid_sync(ctl, nwg))
		return;
	const uint64_t t0 = ctl[C_T0];
	for (;;) {
		/* ---- A: retire, admit, stop? ---- */
		if (blockIdx.x == 0 && tid == 0) {
			uint64_t pooled = xchg(ctl + C_POOLED, (uint64_t)0);
			admit(a, t0, (int64_t)pooled < 0 ? 0u : pooled);
		}
		if (!grid_sync(ctl, nwg))
			return;
		if (ctl[C_STOP])
			break;
		/* ---- B: grow every live job by one block ---- */
		for (uint32_t sl = blockIdx.x; sl < nslices; sl += nwg) {
			const uint32_t s = sl * SV_BLOCK + tid;
			const int64_t job0 = a.slot_job[s];
			const uint32_t steps = job0 >= 0 ? a.slot_steps[s] : 0u;
			/* a slot's block table holds max_len entries; a job never
			 * outgrows its length <= max_len, so this only guards */
			const bool over = job0 >= 0 && steps >= a.max_len;
			c_bad += over;
			const int64_t job = over ? -1 : job0;
			const bool live = job >= 0;
			const uint64_t t = wall_clock64();
			char *p = (char *)kma_wg_malloc<SV_BLOCK, SV_AGG>(H, live ? a.block_bytes : 0u, L);
			const uint64_t dt = wall_clock64() - t;
			const int nl = __syncthreads_count(live);
			const int ns = __syncthreads_count(live && !p);
			if (tid == 0 && nl)
				atomicAdd((unsigned long long *)&a.lat_hist[dt < KMA_SV_NHIST ? dt : KMA_SV_NHIST - 1],
					  (unsigned long long)nl);
			if (tid == 0 && ns)
				fadd(ctl + C_STEP_STALLS, (uint64_t)ns);
			if (live) {
				if (p) {
					a.table[(uint64_t)s * a.max_len + steps] = (uint64_t)(uintptr_t)p;
					*reinterpret_cast<uint64_t *>(p) = pattern(pattern_id(a, (uint64_t)job), steps);
					a.slot_steps[s] = steps + 1u;
					c_mallocs++;
					if (steps + 1u == a.length[job]) {
						uint64_t old = fadd(ctl + C_DONE, ((uint64_t)1 << 40) | (uint64_t)(steps + 1u));
						uint64_t e = old >> 40;
						a.done[4u * e] = s;
						a.done[4u * e + 1u] = (uint64_t)job;
						a.done[4u * e + 2u] = old & MASK40;
						a.done[4u * e + 3u] = steps + 1u;
					}
				} else {
					c_stalls++;
				}
			}
		}
		if (!grid_sync(ctl, nwg))
			return;
		/* ---- C: free the completed jobs' blocks over the whole grid ---- */
		const uint64_t dc = ctl[C_DONE];
		const uint32_t ne = (uint32_t)(dc >> 40);
		const uint64_t T = dc & MASK40;
		for (uint64_t x0 = (uint64_t)blockIdx.x * SV_BLOCK; x0 < T; x0 += (uint64_t)nwg * SV_BLOCK) {
			const uint64_t x = x0 + tid;
			void *q = nullptr;
			if (x < T) {
				uint32_t lo = 0, hi = ne - 1u;
				while (lo < hi) {
					uint32_t mid = (lo + hi + 1u) >> 1;
					if (a.done[4u * mid + 2u] <= x)
						lo = mid;
					else
						hi = mid - 1u;
				}
				const uint64_t sf = a.done[4u * lo], job = a.done[4u * lo + 1u];
				const uint64_t slot = sf & ~PFLAG;
				const uint64_t k = x - a.done[4u * lo + 2u];
				q = (void *)(uintptr_t)a.table[slot * a.max_len + k];
				const uint64_t v = *reinterpret_cast<const uint64_t *>(q);
				c_bad += v != pattern(pattern_id(a, job), k);
				if (!(sf & PFLAG))
					atomicAdd((unsigned long long *)&a.result[job], (unsigned long long)v);
				c_frees++;
			}
			kma_wg_free<SV_BLOCK, SV_AGG>(H, q, L);
		}
		if (tid < 64) {
			uint64_t part = kma_free_superblocks_part(H, blockIdx.x, nwg);
			if (tid == 0)
				fadd(ctl + C_POOLED, part);
		}
		if (!grid_sync(ctl, nwg))
			return;
	}
	/* ---- drain: free the blocks of jobs still live (not completions); the
	 * last phase A retired the previous step's entries and preempted none */
	uint64_t c_drained = 0, c_dfrees = 0;
	for (uint32_t sl = blockIdx.x; sl < nslices; sl += nwg) {
		const uint32_t s = sl * SV_BLOCK + tid;
		const int64_t job = a.slot_job[s];
		uint32_t steps = job >= 0 ? a.slot_steps[s] : 0u;
		/* a job that completed in the last step was retired already */
		c_drained += job >= 0;
		if (tid == 0)
			s_max = 0;
		__syncthreads();
		atomicMax(&s_max, steps);
		__syncthreads();
		const uint32_t m = s_max;
		__syncthreads();
		for (uint32_t k = 0; k < m; k++) {
			void *q = k < steps ? (void *)(uintptr_t)a.table[(uint64_t)s * a.max_len + k] : nullptr;
			c_dfrees += q != nullptr;
			kma_wg_free<SV_BLOCK, SV_AGG>(H, q, L);
		}
		if (job >= 0)
			a.slot_job[s] = -1;
	}
	uint64_t *S = ctl + KMA_SV_STATS;
	const uint64_t v[6] = {c_mallocs, c_stalls, c_frees, c_bad, c_drained, c_dfrees};
	const uint32_t idx[6] = {S_MALLOCS, S_STALLS, S_FREES, S_BAD, S_DRAINED, S_DRAIN_FREES};
	for (int k = 0; k < 6; k++) {
		uint64_t x = v[k];
		for (int o = 32; o > 0; o >>= 1)
			x += __shfl_xor(x, o);
		if (__lane_id() == 0 && x)
			atomicAdd((unsigned long long *)&S[idx[k]], (unsigned long long)x);
	}
}

extern "C" int kmad_serve_max_groups(uint32_t *out)
{
	int dev = 0, per = 0, cus = 0;
	KMA_CHECK(hipGetDevice(&dev));
	KMA_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_serve, SV_BLOCK, 0));
	KMA_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
	*out = (uint32_t)(per * cus);
	return 0;
}

extern "C" int kmad_serve_launch(const kma_serve_args *a, uint32_t nwg, void *stream)
{
	kma_serve_args copy = *a;
	void *params[] = {&copy};
	KMA_CHECK(hipLaunchCooperativeKernel((const void *)k_serve, dim3(nwg), dim3(SV_BLOCK), params, 0,
					     (hipStream_t)stream));
	return 0;
}
