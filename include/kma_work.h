/*
 * kma_work.h - workload entry points of libkma.so (tests and benchmarks).
 *
 * These run the reference's benchmark workloads (kma.cl:424-526,
 * tb_kma.c:222-320) and this library's own traces on the device, with the
 * device timing taken by HIP events on the stream the kernels run on.
 * Plain C ABI; host buffers in, host results out.
 */
#ifndef KMA_WORK_H
#define KMA_WORK_H

#include "kma.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-work-item workloads (one work-item = one device thread). */
#define KMA_WL_FIXED    0u /* malloc(fixed), write {gid, round}, read back, free */
#define KMA_WL_LOGUNI   1u /* malloc(2^e) e in 3..11, hold 0..64 u32 writes, read back, free */
#define KMA_WL_REF_TEST 2u /* kma_test_malloc, kma.cl:424-461 */
#define KMA_WL_LOWVAR   3u /* kma_test_malloc_lowvar, kma.cl:463-492 */
#define KMA_WL_HIGHVAR  4u /* kma_test_malloc_highvar, kma.cl:494-526 */

/* Execution forms. */
#define KMA_FORM_PLAIN   0u /* one kernel launch per round (tb_kma.c:263), per-lane API */
#define KMA_FORM_BATCHED 1u /* one graph-captured persistent launch, all rounds,
			       work-group collective API (native layout only) */
#define KMA_FORM_FUSED   2u /* one launch, all rounds, per-lane API */

typedef struct kma_workload {
	uint32_t mode;    /* KMA_WL_* */
	uint32_t form;    /* KMA_FORM_* */
	uint64_t nwi;     /* work-items */
	uint32_t block;   /* work-group size (multiple of 64, <= 256) */
	uint32_t rounds;  /* rounds (reference: `iters`) */
	uint32_t round0;  /* first round index (trace position) */
	uint32_t fixed;   /* size for KMA_WL_FIXED */
	uint64_t seed;
	uint32_t warmup;  /* untimed rounds run first */
	uint32_t reps;    /* reserved, 0 */
	uint32_t variant; /* KMA_VAR_* bits */
} kma_workload;

/* Workload variants. */
#define KMA_VAR_LANE_MALLOC 0x1u   /* BATCHED: per-lane kma_malloc instead of kma_wg_malloc */
#define KMA_VAR_LANE_FREE   0x2u   /* BATCHED: per-lane kma_free instead of kma_wg_free */
#define KMA_VAR_LANE_CLAIM  0x4u   /* BATCHED: one bitmap atomic per block, not per word */
#define KMA_VAR_LATENCY     0x100u /* record every malloc's latency (wall clock) */
#define KMA_VAR_PHASES      0x200u /* time the malloc / hold / free phases */

typedef struct kma_run_stats {
	double gpu_ms;          /* HIP-event time of the timed region */
	double kernel_ms_avg;   /* average duration of one launch in it */
	uint64_t launches;      /* launches in the timed region */
	uint64_t ops;           /* mallocs + frees issued in the timed region */
	uint64_t mallocs;
	uint64_t nulls;         /* mallocs that returned NULL */
	uint64_t bad;           /* pattern read-back mismatches */
	uint64_t double_frees;  /* heap counter (native layout) */
	uint64_t bad_frees;     /* heap counter (native layout) */
	uint64_t pool_empty;    /* refills that found no free superblock */
	uint64_t waits[4];      /* watchdog expiries: reserve, claim, wavefront claim, pool */
	uint64_t remote_frees;  /* frees issued from a non-home XCD */
	double lat_p50_ns;      /* KMA_VAR_LATENCY: malloc latency percentiles */
	double lat_p99_ns;
	double lat_max_ns;
	double phase_ns[3];     /* KMA_VAR_PHASES: malloc / hold / free time per
				   wavefront-round */
} kma_run_stats;

/* Run a per-work-item workload on the heap. */
int kma_run_workload(kma_heap *h, const kma_workload *wl, kma_run_stats *out);

/* Replay an op list on ONE device work-item.  ops[i] >= 0: malloc(ops[i]),
 * out[i] = heap offset of the block (0 = NULL); ops[i] < 0: free the block
 * of op -ops[i]-1.  pat (optional) is written to each block's first word. */
int kma_trace_serial(kma_heap *h, const int64_t *ops, uint64_t n, uint64_t *out,
		     const uint32_t *pat);

/* Agent-scope atomicOr throughput over a `bytes`-byte device buffer:
 * pattern 0 distinct pseudo-random words, 1 coalesced consecutive words,
 * 2 one word per wavefront.  nthreads lanes x iters atomics; the first
 * `warmup` launches are untimed.  Returns atomics per second in *ops_per_s
 * and the average launch time in *ms. */
int kma_atomic_peak(uint64_t bytes, uint64_t nthreads, uint32_t iters, int pattern,
		    uint32_t warmup, uint32_t reps, double *ops_per_s, double *ms);

/* Self-check of the wavefront bitmap claim on a native heap (test hook):
 * case i claims got[i] blocks of class cls[i] from superblock i, whose
 * bitmap (bm_words words) is first set to bm_in[i]; returns the block
 * offsets in res_out[i*64 ..], the final bitmap in bm_out and the number
 * claimed in done_out[i].  Destroys the heap's contents. */
int kma_check_wave_claim(kma_heap *h, uint32_t ncases, const uint32_t *cls, const uint32_t *base,
			 const uint32_t *got, const uint32_t *bm_in, uint64_t *res_out,
			 uint32_t *bm_out, uint32_t *done_out);

/* Native layout: detach every context's superblock so that fully free
 * superblocks return to the pool (no allocation may be in flight). */
int kma_trim(kma_heap *h);

#ifdef __cplusplus
}
#endif
#endif
