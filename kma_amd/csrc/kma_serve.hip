/*
 * kma_serve.hip - the device-resident serving loop (kma_serve_run in
 * include/kma_work.h): one cooperative persistent kernel that admits jobs as
 * they arrive, grows every live job by one block per step and frees a job's
 * blocks when it completes, on a native heap.
 *
 * One step = three phases separated by grid barriers:
 *   A (the 256 threads of work-group 0): retire the jobs completed in the
 *     previous step, one done-list entry per thread (slot back to the
 *     free-slot stack at an LDS-counted position); then admit arrived jobs in
 *     arrival order (requeued jobs first) while a slot is free and the
 *     admission policy allows; decide whether the run is over;
 *   B (all work-groups): work-group g serves job slots [256 k, 256 k + 256)
 *     for k = (g + step) mod S, (g + G + step) mod S, ... (S slices of 256
 *     slots; the rotation takes every job's blocks from every XCD's
 *     superblocks alike): ONE collective kma_wg_malloc of block_bytes for
 *     every live slot, the block recorded in the slot's block table and
 *     stamped with pattern(job, step); a job that reaches its length appends
 *     (slot, job, first, count) to the done list, whose 64-bit counter keeps
 *     the entry index in bits 40.. and the running block total below, so
 *     entries are ordered by `first`;
 *   C (all work-groups): the blocks of all completed jobs, numbered
 *     0 .. total-1, are spread over the whole grid: each thread finds its
 *     entry by binary search, checks the block's pattern, adds it to the
 *     job's result and frees the block with kma_wg_free.
 *
 * Projected admission.  A job admitted at step s with length L holds
 * k + 1 blocks after phase B of step s + k (k < L) and none after phase C of
 * step s + L - 1, so the live set's use after every future step is known.
 * Work-group 0 keeps it in LDS as a ring of KMA_SERVE_MAX_PROJECTED_LEN u32
 * entries, entry t % len holding the projected blocks in use after phase B
 * of step t (valid for t in [s, s + len)).  Admitting a job checks, one step
 * ahead per thread, that entry + k + 1 stays within cap_blocks for all
 * k < L (one __syncthreads_or), then adds the ramp; the entry of step s - 1
 * is cleared when step s starts, since no ramp reaches that far ahead.
 * Completions need no update: a ramp ends by itself.  With every live job's
 * future use inside the heap, a NULL can only come from blocks held in
 * other contexts' partly used superblocks (the watermark's headroom covers
 * them); if one comes anyway, the youngest jobs are preempted as below and
 * the timeline is rebuilt from the live slots (slot steps and lengths).
 *
 * Watermark admission (the policy of a use-watermark serving system) admits
 * while the superblocks out of the pool stay below the watermark.  It cannot
 * see each job's future growth, so a full heap is met with NULLs: after a
 * step with stalls, phase A preempts the most recently admitted live jobs
 * until their blocks cover the shortfall (their blocks are freed in phase C,
 * without adding to their results) and puts them on a requeue ring that is
 * admitted before new arrivals; admission then pauses until use falls
 * SV_HYST_PCT points below the watermark.
 *
 * Barriers release (write back L2) and acquire (invalidate) at agent scope,
 * so plain stores of one phase are visible to every XCD in the next.  Each
 * barrier wait has a 5 s watchdog that aborts the whole loop.  Work-group 0
 * reads the wall clock at every phase boundary (phase profile in the stats),
 * and every work-group adds up the time it spends working in phases B and C.
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
	C_NFREE = 48, /* free slots (phase A keeps the stack top in LDS) */
	C_STOP = 64,
	C_DONE = 80, /* done list: entries << 40 | blocks */
	C_T0 = 96,
	C_ABORT = 112,
	C_LIVE = 128,
	C_MTOT = 144,      /* successful mallocs so far (phase B adds per work-group) */
	C_LTOP = 192,      /* phase A only: live-stack top, requeue head/tail, sequence */
	C_RQ_HEAD = 193,
	C_RQ_TAIL = 194,
	C_SEQ = 195,
	C_HOLD = 196,      /* watermark admission paused since a step with stalls */
	C_FPREV = 197,     /* C_FTOT as of the previous phase A */
	C_REBUILT = 198,   /* the previous step rebuilt the timeline */
	C_FTOT = 208,      /* blocks freed so far in phase C (added per work-group) */
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
	S_REBUILDS,
	S_PEAK_PROJ,
	S_LIVE_SUM,
	S_PH_AWORK, /* work-group 0's wall clock per phase, summed over the steps */
	S_PH_A,
	S_PH_B,
	S_PH_C,
	S_WORK_B, /* every work-group's working time in phases B and C, summed */
	S_WORK_C,
	S_PROJ_CHECKED, /* steps whose projected use was compared with the blocks held */
	S_PROJ_MISMATCH,
	S_N
};
constexpr uint64_t MASK40 = (1ull << 40) - 1ull;
constexpr uint32_t SV_HYST_PCT = 10; /* admission hysteresis, points below the watermark */
constexpr uint64_t BARRIER_TICKS = 500000000ull; /* 5 s */
constexpr uint32_t SV_TL = KMA_SERVE_MAX_PROJECTED_LEN;
static_assert((SV_TL & (SV_TL - 1u)) == 0, "timeline length is a power of two");

/* Work-group 0's LDS for phase A: the projected-use timeline and the values
 * its threads share. */
struct sv_admit_lds {
	uint32_t tl[SV_TL];
	uint32_t part[SV_BLOCK];       /* ring_scan: per-thread chunk totals */
	uint32_t nfree;                /* free-slot stack top (a.free_slots) */
	uint32_t n_free, n_rq, n_done; /* retirement: positions counted in LDS */
	uint32_t rebuild, cand_len;
	int64_t cand;
	uint64_t now, step;
};

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

/* Inclusive prefix sum (mod 2^32) of the timeline ring in step order from
 * entry `start`: thread t sums its SV_TL / SV_BLOCK consecutive entries,
 * the chunk totals are scanned, then added back.  All threads call it. */
__device__ __forceinline__ void ring_scan(sv_admit_lds &P, uint64_t start)
{
	constexpr uint32_t PER = SV_TL / SV_BLOCK;
	const uint32_t tid = threadIdx.x;
	uint32_t acc = 0;
	for (uint32_t j = 0; j < PER; j++) {
		uint32_t &e = P.tl[(start + tid * PER + j) & (SV_TL - 1u)];
		acc += e;
		e = acc;
	}
	P.part[tid] = acc;
	__syncthreads();
	if (tid == 0) {
		uint32_t run = 0;
		for (uint32_t i = 0; i < SV_BLOCK; i++) {
			const uint32_t v = P.part[i];
			P.part[i] = run;
			run += v;
		}
	}
	__syncthreads();
	const uint32_t off = P.part[tid];
	for (uint32_t j = 0; j < PER; j++)
		P.tl[(start + tid * PER + j) & (SV_TL - 1u)] += off;
	__syncthreads();
}

/* Work-group 0, all threads, once: the free-slot stack holds every slot,
 * ordered so that successive pops cycle over the slices (pop k: slice
 * k mod S, lane k / S), which spreads the live jobs evenly over the slices
 * and so over the work-groups that grow them; a retired slot goes back on
 * top, so the next admission refills the same slice. */
__device__ inline void init_slot_stack(const kma_serve_args &a, sv_admit_lds &P)
{
	const uint32_t S = a.max_live / SV_BLOCK;
	for (uint32_t k = threadIdx.x; k < a.max_live; k += SV_BLOCK)
		a.free_slots[a.max_live - 1u - k] = (k % S) * SV_BLOCK + k / S;
	if (threadIdx.x == 0)
		P.nfree = a.max_live;
	__syncthreads();
}

/* Phase A: all SV_BLOCK threads of work-group 0 (P is that group's LDS).
 * Thread 0 keeps the loop's scalar state; the other threads retire done-list
 * entries and check and add timeline ramps with it. */
__device__ __forceinline__ void phase_a(const kma_serve_args &a, uint64_t t0, uint64_t pooled,
					sv_admit_lds &P)
{
	const kma_dheap *H = (const kma_dheap *)a.heap;
	const uint32_t tid = threadIdx.x;
	const bool projected = a.admit == KMA_SERVE_ADMIT_PROJECTED;
	uint64_t *ctl = a.ctl;
	uint64_t *S = ctl + KMA_SV_STATS;
	/* thread 0's state (the other threads' copies are unused) */
	uint64_t live = 0, next = 0, ltop = 0, rqh = 0, rqt = 0, seq = 0, dc = 0, used = 0;
	bool stop = false, hold = false;
	if (tid == 0) {
		live = ctl[C_LIVE];
		next = ctl[C_NEXT];
		ltop = ctl[C_LTOP];
		rqh = ctl[C_RQ_HEAD];
		rqt = ctl[C_RQ_TAIL];
		seq = ctl[C_SEQ];
		dc = ctl[C_DONE];
		P.now = wall_clock64() - t0;
		P.step = S[S_STEPS];
		P.n_free = P.n_rq = P.n_done = 0;
	}
	__syncthreads();
	const uint64_t now = P.now, step = P.step;
	/* ---- retire the previous step's entries: completions and preemptions,
	 * one entry per thread; the slot goes back on the free-slot stack ---- */
	{
		const uint32_t ne = (uint32_t)(ctl[C_DONE] >> 40);
		const uint64_t rqt0 = ctl[C_RQ_TAIL];
		for (uint32_t e = tid; e < ne; e += SV_BLOCK) {
			const uint64_t sf = a.done[4u * e], job = a.done[4u * e + 1u];
			const uint32_t slot = (uint32_t)(sf & ~PFLAG);
			if (sf & PFLAG) {
				a.requeue[(rqt0 + atomicAdd(&P.n_rq, 1u)) % a.rq_cap] = job;
			} else {
				a.job_done[job] = now + 1u;
				a.slot_job[slot] = -1;
				atomicAdd(&P.n_done, 1u);
			}
			a.free_slots[P.nfree + atomicAdd(&P.n_free, 1u)] = slot;
		}
	}
	__syncthreads();
	if (tid == 0) {
		P.nfree += P.n_free;
		rqt += P.n_rq;
		live -= P.n_done;
		S[S_DONE] += P.n_done;
		dc = 0;
		const uint64_t stalls = ctl[C_STEP_STALLS];
		ctl[C_STEP_STALLS] = 0;
		used = H->nsb > pooled ? H->nsb - pooled : 0u;
		if (used > S[S_PEAK_USED])
			S[S_PEAK_USED] = used;
		stop = now >= a.duration_ticks || (next >= a.njobs && live == 0 && rqh == rqt);
		if (projected) {
			hold = stalls != 0; /* this step: rebuild, admit nothing */
		} else {
			/* hysteresis: a step with stalls pauses admission until use
			 * falls SV_HYST_PCT points below the watermark, so the heap is
			 * not refilled with jobs that the next steps' growth would
			 * preempt again */
			hold = ctl[C_HOLD] != 0;
			if (stalls)
				hold = true;
			else if (hold && used * 100u < (uint64_t)(a.watermark_pct > SV_HYST_PCT ? a.watermark_pct - SV_HYST_PCT
												    : 0u) *
							    H->nsb)
				hold = false;
			ctl[C_HOLD] = hold ? 1u : 0u;
		}
		if (projected && step > 0u) {
			/* the previous step's projection against the blocks it held
			 * after its growth phase: successful mallocs so far minus the
			 * blocks freed before that phase; equal unless the step stalled
			 * or held the blocks of jobs preempted at its start */
			const uint64_t held = ctl[C_MTOT] - ctl[C_FPREV];
			if (!stalls && !ctl[C_REBUILT]) {
				S[S_PROJ_CHECKED]++;
				S[S_PROJ_MISMATCH] += held != P.tl[(step - 1u) & (SV_TL - 1u)];
			}
		}
		ctl[C_FPREV] = ctl[C_FTOT];
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
		P.rebuild = projected && !stop && stalls;
		ctl[C_REBUILT] = P.rebuild;
		if (projected && !P.rebuild && step > 0u)
			P.tl[(step - 1u) & (SV_TL - 1u)] = 0u; /* now stands for step - 1 + SV_TL */
	}
	__syncthreads();
	if (P.rebuild) {
		/* the timeline from the live slots: a job holding c blocks with
		 * r = L - c >= 1 steps to go holds v_k = c + k + 1 after step + k,
		 * k < r, and 0 from k = r on.  Its second differences are c + 1 at
		 * k = 0, -c at 1, -(c + r) - 1 at r and c + r at r + 1 (terms at
		 * SV_TL or beyond only act past the window): two LDS atomics per
		 * job, then two prefix sums in step order give the sum of the v */
		for (uint32_t i = tid; i < SV_TL; i += SV_BLOCK)
			P.tl[i] = 0u;
		__syncthreads();
		for (uint32_t s = tid; s < a.max_live; s += SV_BLOCK) {
			const int64_t job = a.slot_job[s];
			if (job < 0)
				continue;
			const uint32_t c = a.slot_steps[s], r = a.length[job] - c;
			atomicAdd(&P.tl[step & (SV_TL - 1u)], c + 1u);
			atomicAdd(&P.tl[(step + 1u) & (SV_TL - 1u)], 0u - c);
			if (r < SV_TL)
				atomicAdd(&P.tl[(step + r) & (SV_TL - 1u)], 0u - (c + r + 1u));
			if (r + 1u < SV_TL)
				atomicAdd(&P.tl[(step + r + 1u) & (SV_TL - 1u)], c + r);
		}
		__syncthreads();
		ring_scan(P, step);
		ring_scan(P, step);
	}
	/* ---- admission, oldest first ---- */
	uint32_t nadm = 0;
	if (projected) {
		const uint32_t cap = (uint32_t)a.cap_blocks;
		for (;;) {
			if (tid == 0) {
				int64_t c = -1;
				if (!stop && !hold && P.nfree > 0 && nadm < a.max_admit) {
					if (rqh != rqt)
						c = (int64_t)a.requeue[rqh % a.rq_cap];
					else if (next < a.njobs && a.arrival[next] <= now)
						c = (int64_t)next;
				}
				P.cand = c;
				P.cand_len = c >= 0 ? a.length[c] : 0u;
			}
			__syncthreads();
			const int64_t job = P.cand;
			if (job < 0)
				break;
			const uint32_t len = P.cand_len;
			int over = 0;
			for (uint32_t k = tid; k < len; k += SV_BLOCK)
				over |= P.tl[(step + k) & (SV_TL - 1u)] + k + 1u > cap;
			if (__syncthreads_or(over))
				break; /* first come, first served: later arrivals wait too */
			for (uint32_t k = tid; k < len; k += SV_BLOCK)
				P.tl[(step + k) & (SV_TL - 1u)] += k + 1u;
			if (tid == 0) {
				if (rqh != rqt) {
					rqh++;
				} else {
					next++;
					a.job_admit[job] = now + 1u;
				}
				const uint32_t slot = a.free_slots[--P.nfree];
				a.slot_job[slot] = job;
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
				nadm++;
			}
		}
	} else if (tid == 0 && !stop && !hold) {
		while (P.nfree > 0 && nadm < a.max_admit && used * 100u < (uint64_t)a.watermark_pct * H->nsb) {
			uint64_t job;
			if (rqh != rqt) {
				job = a.requeue[rqh++ % a.rq_cap];
			} else if (next < a.njobs && a.arrival[next] <= now) {
				job = next++;
				a.job_admit[job] = now + 1u;
			} else {
				break;
			}
			const uint32_t slot = a.free_slots[--P.nfree];
			a.slot_job[slot] = (int64_t)job;
			a.slot_steps[slot] = 0;
			a.slot_seq[slot] = (uint32_t)++seq;
			if (ltop == a.lstack_cap) {
				uint64_t w = 0;
				for (uint64_t r = 0; r < ltop; r++)
					if (lentry_live(a, a.lstack[r]))
						a.lstack[w++] = a.lstack[r];
				ltop = w;
			}
			a.lstack[ltop++] = (seq << 32) | slot;
			live++;
			nadm++;
		}
	}
	if (tid == 0) {
		S[S_ADMITTED] += nadm;
		if (projected) {
			const uint64_t proj = P.tl[step & (SV_TL - 1u)];
			if (proj > S[S_PEAK_PROJ])
				S[S_PEAK_PROJ] = proj;
		}
		if (!stop) {
			S[S_STEPS]++;
			S[S_LIVE_SUM] += live;
			S[S_REBUILDS] += P.rebuild;
		} else {
			S[S_RUN_TICKS] = now;
		}
		ctl[C_NFREE] = P.nfree;
		ctl[C_LIVE] = live;
		ctl[C_NEXT] = next;
		ctl[C_LTOP] = ltop;
		ctl[C_RQ_HEAD] = rqh;
		ctl[C_RQ_TAIL] = rqt;
		ctl[C_SEQ] = seq;
		ctl[C_STOP] = stop ? 1u : 0u;
	}
}

} // namespace

__global__ void __launch_bounds__(SV_BLOCK) k_serve(kma_serve_args a)
{
	using namespace kma_nat;
	kma_dheap *H = (kma_dheap *)a.heap;
	__shared__ kma_wg_lds<SV_BLOCK> L;
	__shared__ sv_admit_lds P; /* used by work-group 0 only */
	__shared__ uint32_t s_max;
	kma_wg_lds_init(L);
	const uint32_t tid = threadIdx.x, nwg = gridDim.x;
	const bool wg0 = blockIdx.x == 0;
	uint64_t *ctl = a.ctl;
	const uint32_t nslices = a.max_live / SV_BLOCK;
	uint64_t c_mallocs = 0, c_stalls = 0, c_frees = 0, c_bad = 0;
	/* phase profile: work-group 0's wall clock per phase (thread 0) and
	 * every group's working time in phases B and C */
	uint64_t ph_awork = 0, ph_a = 0, ph_b = 0, ph_c = 0, w_b = 0, w_c = 0;
	if (wg0 && tid == 0)
		ctl[C_T0] = wall_clock64();
	if (wg0)
		for (uint32_t i = tid; i < SV_TL; i += SV_BLOCK)
			P.tl[i] = 0u;
	/* every work-group counts its slice of the heap's free superblocks
	 * (here, then after each step's frees); phase A reads the sum */
	if (tid < 64) {
		uint64_t part = kma_free_superblocks_part(H, blockIdx.x, nwg);
		if (tid == 0)
			fadd(ctl + C_POOLED, part);
	}
	if (!grid_sync(ctl, nwg))
		return;
	if (wg0)
		init_slot_stack(a, P);
	const uint64_t t0 = ctl[C_T0];
	uint64_t tA = wall_clock64();
	for (uint64_t lstep = 0;; lstep++) {
		/* ---- A: retire, admit, stop? ---- */
		if (wg0) {
			__shared__ uint64_t s_pooled;
			if (tid == 0) {
				uint64_t pooled = xchg(ctl + C_POOLED, (uint64_t)0);
				s_pooled = (int64_t)pooled < 0 ? 0u : pooled;
			}
			__syncthreads();
			phase_a(a, t0, s_pooled, P);
			if (tid == 0)
				ph_awork += wall_clock64() - tA;
		}
		if (!grid_sync(ctl, nwg))
			return;
		const uint64_t tB = wall_clock64();
		ph_a += tB - tA;
		if (ctl[C_STOP])
			break;
		/* ---- B: grow every live job by one block.  Work-group j grows
		 * slice (j + step) mod nslices (and j + nwg, ...): a slice moves to
		 * the next work-group, hence to the next XCD, every step, so every
		 * job takes its blocks from all XCDs' superblocks alike ---- */
		for (uint32_t j = blockIdx.x; j < nslices; j += nwg) {
			const uint32_t sl = (uint32_t)((j + lstep) % nslices);
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
			if (tid == 0 && nl > ns)
				fadd(ctl + C_MTOT, (uint64_t)(nl - ns));
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
		w_b += wall_clock64() - tB;
		if (!grid_sync(ctl, nwg))
			return;
		const uint64_t tC = wall_clock64();
		ph_b += tC - tB;
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
		if (tid == 0 && (uint64_t)blockIdx.x * SV_BLOCK < T) {
			/* blocks this group freed: min(256, T - x0) per x0 */
			const uint64_t stride = (uint64_t)nwg * SV_BLOCK, x0 = (uint64_t)blockIdx.x * SV_BLOCK;
			const uint64_t full = (T - x0) / stride, rem = (T - x0) % stride;
			fadd(ctl + C_FTOT, full * SV_BLOCK + (rem < SV_BLOCK ? rem : SV_BLOCK));
		}
		if (tid < 64) {
			uint64_t part = kma_free_superblocks_part(H, blockIdx.x, nwg);
			if (tid == 0)
				fadd(ctl + C_POOLED, part);
		}
		w_c += wall_clock64() - tC;
		if (!grid_sync(ctl, nwg))
			return;
		tA = wall_clock64();
		ph_c += tA - tC;
	}
	/* ---- drain: free the blocks of jobs still live (not completions); the
	 * last phase A retired the previous step's entries and preempted none */
	uint64_t c_drained = 0, c_dfrees = 0;
	for (uint32_t sl = blockIdx.x; sl < (a.keep_live ? 0u : nslices); sl += nwg) {
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
	if (tid == 0) {
		if (wg0) {
			S[S_PH_AWORK] = ph_awork;
			S[S_PH_A] = ph_a;
			S[S_PH_B] = ph_b;
			S[S_PH_C] = ph_c;
		}
		atomicAdd((unsigned long long *)&S[S_WORK_B], (unsigned long long)w_b);
		atomicAdd((unsigned long long *)&S[S_WORK_C], (unsigned long long)w_c);
	}
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
