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
	uint32_t deadline_ms; /* reference layout: every allocator wait that may
				 fail gives up (NULL, counted in deadline_hits)
				 once this many ms of the timed region have passed;
				 0: no deadline, only the 4 s per-call watchdogs */
	uint32_t variant; /* KMA_VAR_* bits */
} kma_workload;

/* Workload variants. */
#define KMA_VAR_LANE_MALLOC 0x1u   /* BATCHED: per-lane kma_malloc instead of kma_wg_malloc */
#define KMA_VAR_LANE_FREE   0x2u   /* BATCHED: per-lane kma_free instead of kma_wg_free */
#define KMA_VAR_LANE_CLAIM  0x4u   /* BATCHED: one bitmap atomic per block, not per word */
/* BATCHED forms (workloads and streams): aggregation width of the
 * collective calls (kma_wg_device.h).  WAVE: each wavefront serves its
 * lanes (no LDS, no barrier); WG8/16/32: the work-group stages its
 * requests in LDS and claims with 8-, 16- or 32-lane groups. */
#define KMA_VAR_AGG_SHIFT   4u
#define KMA_VAR_AGG_MASK    0x30u
#define KMA_VAR_AGG_WAVE    0x00u
#define KMA_VAR_AGG_WG8     0x10u
#define KMA_VAR_AGG_WG16    0x20u
#define KMA_VAR_AGG_WG32    0x30u
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
	uint64_t waits[4];      /* native layout: watchdog expiries of reserve, claim,
				   group claim, pool.  Reference layout: give-ups of
				   reserve (watchdog: NULL), claim (slot returned: NULL),
				   dequeue (refill saw an empty queue) and enqueue
				   (corrupt queue: superblock lost) */
	uint64_t remote_frees;  /* frees issued from a non-home XCD */
	uint64_t migrated;      /* superblocks moved from an XCD's dirty shards to
				   its clean shards inside a kernel */
	double lat_p50_ns;      /* KMA_VAR_LATENCY: malloc latency percentiles */
	double lat_p99_ns;
	double lat_max_ns;
	double phase_ns[3];     /* KMA_VAR_PHASES: malloc / hold / free time per
				   wavefront-round */
	uint64_t deadline_hits; /* reference layout: mallocs that returned NULL
				   because deadline_ms had passed */
	uint64_t listed;        /* native: detached superblocks put into a partial index */
	uint64_t reattached;    /* native: listed superblocks reused (attached to a head again, or grabbed) */
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
 * 2 one word per wavefront; pattern 3 instead issues plain 4-B stores,
 * `iters` consecutive words per lane into the lane's own 2 KiB slot (the
 * store shape of the full-occupancy workload).  nthreads lanes x iters
 * operations; the first `warmup` launches are untimed.  Returns operations
 * per second in *ops_per_s and the average launch time in *ms. */
int kma_atomic_peak(uint64_t bytes, uint64_t nthreads, uint32_t iters, int pattern,
		    uint32_t warmup, uint32_t reps, double *ops_per_s, double *ms);

/* Streaming copy of `bytes` bytes between two device buffers (k_copy):
 * *gbps = (read + written bytes) per second, *ms = average launch time.
 * The HBM bandwidth reference, and the known-size kernel that scales the
 * FETCH_SIZE / WRITE_SIZE counters. */
int kma_copy_probe(uint64_t bytes, uint32_t warmup, uint32_t reps, double *gbps, double *ms);

/* Self-check of the bitmap claim of the work-group collective malloc (one
 * 16-lane group per case, four cases per wavefront) on a native heap (test
 * hook): case i claims got[i] blocks of class cls[i] from superblock i, whose
 * bitmap (bm_words words) is first set to bm_in[i]; returns the block
 * offsets in res_out[i*64 ..], the final bitmap in bm_out and the number
 * claimed in done_out[i].  Destroys the heap's contents. */
int kma_check_wave_claim(kma_heap *h, uint32_t ncases, const uint32_t *cls, const uint32_t *base,
			 const uint32_t *got, const uint32_t *bm_in, uint64_t *res_out,
			 uint32_t *bm_out, uint32_t *done_out);

/* Test hook with no counterpart in the reference: ONE wavefront frees the
 * blocks at heap offsets offs[0..n-1] (n <= 64; lane i frees offs[i], 0 = no
 * request) in a single kma_free call, so that lanes freeing neighbouring
 * slots of one bitmap word are combined as in a kernel.  Native layout; the
 * offsets must lie in the heap's data (block offsets as kma_trace_serial
 * returns them); double frees are counted in the heap's statistics. */
int kma_wave_free(kma_heap *h, const uint64_t *offs, uint32_t n);

/* Native layout: detach every context's superblock so that fully free
 * superblocks return to the pool (no allocation may be in flight). */
int kma_trim(kma_heap *h);

/* ------------------------------------------------------------------ */
/* Batched request streams                                              */
/* ------------------------------------------------------------------ */
/* A stream is a host-generated trace of allocation batches: batch sizes are
 * Zipf-distributed, each request is either a malloc of a log-uniform size or
 * (with probability free_pct %) a free of a block returned by an EARLIER
 * batch, chosen uniformly among the blocks still live.  With `drain` a last
 * batch frees every block still live, so a run returns the heap to empty.
 * Request j is encoded in one u32: bit 31 clear: malloc(j's low 31 bits);
 * bit 31 set: free the block of request (low 31 bits). */
typedef struct kma_stream_spec {
	uint32_t nbatch;     /* batches before the drain batch */
	uint32_t max_batch;  /* batch sizes Zipf over 1..max_batch */
	uint32_t zipf_milli; /* Zipf exponent x 1000 */
	uint32_t free_pct;   /* percent of requests that are frees */
	uint32_t min_size;   /* log-uniform sizes in [min_size, max_size), powers of two */
	uint32_t max_size;
	uint32_t page_min;   /* sizes >= page_min are counted as large in the info */
	uint32_t drain;      /* 1: append the drain batch */
	uint64_t seed;
} kma_stream_spec;

typedef struct kma_stream_info {
	uint64_t nreq;        /* requests, drain included */
	uint64_t nbatch;      /* batches, drain included */
	uint64_t nmalloc;
	uint64_t nfree;
	uint64_t max_batch;   /* largest batch drawn */
	uint64_t peak_live;   /* most blocks live at a batch boundary */
	uint64_t small_bytes; /* sum over mallocs < page_min of the size rounded up to a power of two */
	uint64_t large_bytes; /* sum over mallocs >= page_min of the size rounded up to page_min multiples */
	uint64_t peak_small_bytes; /* same sums over the blocks live at the worst batch boundary */
	uint64_t peak_large_bytes;
} kma_stream_info;

typedef struct kma_stream kma_stream; /* opaque: host trace + device copy */

/* Generate the trace on the host and upload it.  NULL on bad input. */
kma_stream *kma_stream_create(const kma_stream_spec *spec);
void kma_stream_destroy(kma_stream *s);
int kma_stream_get_info(const kma_stream *s, kma_stream_info *out);
/* Copy the trace: req[nreq], batch_off[nbatch + 1]. */
int kma_stream_trace(const kma_stream *s, uint32_t *req, uint32_t *batch_off);
/* After a run: per request, the heap offset a malloc returned (0: NULL;
 * 0 for frees). */
int kma_stream_results(const kma_stream *s, uint64_t *res);
/* After a run with KMA_VAR_LOG: per request, the device wall-clock tick at
 * which a malloc returned or a free was issued. */
int kma_stream_times(const kma_stream *s, uint64_t *ticks);

#define KMA_VAR_LOG 0x400u /* record per-request wall-clock ticks */

typedef struct kma_stream_stats {
	double gpu_ms;         /* HIP-event time of one replay (average over reps) */
	double kernel_ms_avg;  /* average launch duration inside it */
	uint64_t launches;     /* launches per replay */
	uint64_t ops;          /* requests executed (mallocs + frees) per replay */
	uint64_t mallocs;      /* mallocs that returned a block */
	uint64_t nulls;        /* mallocs that returned NULL */
	uint64_t frees;        /* frees of a block */
	uint64_t free_nulls;   /* frees whose malloc had returned NULL (no-ops) */
	uint64_t bad;          /* blocks whose first word lost the pattern written at malloc */
	uint64_t wait_timeouts;/* frees that gave up waiting for their malloc's result */
	uint64_t large;        /* mallocs above the largest class (pages) */
	uint64_t large_nulls;
	uint64_t double_frees, bad_frees, pool_empty, waits[4], remote_frees; /* heap counters, per replay */
	double tick_ns;        /* wall-clock tick length for kma_stream_times */
	double phase_ns[3];    /* KMA_VAR_PHASES: malloc / free-target wait / free time per
				  wavefront slice (plain: per wavefront launch) */
	uint64_t listed, reattached; /* heap counters, per replay: partial-index puts and re-attachments */
} kma_stream_stats;

/* Run the stream on a native heap.
 * KMA_FORM_PLAIN:   one launch per batch, per-lane kma_malloc / kma_free.
 * KMA_FORM_BATCHED: ONE graph-captured persistent launch of nwg work-groups
 *   (256 threads) that claim chunks of 256 * per_thread consecutive requests
 *   and run the chunk's 256-request slices through kma_wg_malloc, then
 *   through kma_wg_free, at the aggregation width variant & KMA_VAR_AGG_MASK
 *   selects; a free waits for the result of its malloc (always an earlier
 *   request).
 * warmup untimed replays, then reps timed replays; the heap must be empty
 * before each replay (true after a draining replay). */
int kma_stream_run(kma_heap *h, kma_stream *s, uint32_t form, uint32_t nwg, uint32_t per_thread,
		   uint32_t warmup, uint32_t reps, uint32_t variant, kma_stream_stats *out);

/* Heap configuration for a stream: `heap_bytes` split between superblocks
 * and pages in proportion to the stream's small / large demand
 * (kma_stream_info), superblocks of 2^sb_log2 bytes, class table `table`:
 *   KMA_TABLE_POW2: powers of two from min_size up to min(SB, page_min);
 *   KMA_TABLE_HALF: the same plus the midpoints 1.5 x 2^k.
 * Requests above the largest class take 2^page_log2-byte pages. */
#define KMA_TABLE_POW2 0u
#define KMA_TABLE_HALF 1u
int kma_stream_heap_config(const kma_stream *s, uint64_t heap_bytes, uint32_t sb_log2, uint32_t table,
			   uint32_t page_log2, kma_config *out);

#define KMA_TUNE_MAX 128
typedef struct kma_tune_trial {
	uint32_t sb_log2, table, nwg, per_thread;
	uint32_t agg;         /* aggregation width: a KMA_VAR_AGG_* value */
	uint32_t no_partial;  /* the heap's kma_config.no_partial */
	double ops_per_s;     /* requests per second */
	double ms;            /* per replay */
	uint64_t nulls;
	uint64_t violations;  /* bad + wait_timeouts + double/bad frees + watchdog expiries */
} kma_tune_trial;

typedef struct kma_tune_result {
	kma_config cfg;       /* chosen heap configuration */
	uint32_t nwg, per_thread; /* chosen persistent launch shape */
	uint32_t table;
	uint32_t ntrials;
	uint32_t agg;         /* chosen aggregation width (KMA_VAR_AGG_*): pass it
				 in kma_stream_run's variant */
	double ops_per_s;
	uint64_t nulls;
	kma_tune_trial trials[KMA_TUNE_MAX];
} kma_tune_result;

/* Autotuner for the persistent form: sweeps superblock size (16, 32,
 * 64 KiB), class table (KMA_TABLE_*), aggregation width (wavefront, or
 * work-group staging with 8-, 16- or 32-lane claim groups), work-groups
 * (512, 1024) and requests per thread per chunk (1, 2): 96 trials of `reps`
 * timed replays each, and picks the highest throughput among trials with no
 * violation and at most `null_limit` NULL returns (no limit when 0); then
 * one more trial of the pick on a heap with no_partial set, kept if it is
 * faster under the same conditions (97 trials).  Every trial heap is
 * `heap_bytes`. */
int kma_stream_autotune(kma_stream *s, uint64_t heap_bytes, uint32_t page_log2, uint64_t null_limit,
			uint32_t reps, kma_tune_result *out);

/* ------------------------------------------------------------------ */
/* Serving loop                                                         */
/* ------------------------------------------------------------------ */
/* Jobs arrive as a Poisson process and run for a log-normal number of
 * steps; each step appends one block of block_bytes to every live job, and
 * a job's blocks are all freed when it completes.  One device-resident
 * persistent kernel admits arrived jobs (oldest first) while a job slot is
 * free and the admission policy allows, grows every live job by one block
 * per step with one work-group collective malloc per 256 slots, and frees
 * completed jobs' blocks with the whole grid.
 *
 * Admission policies (spec.admit):
 *   KMA_SERVE_ADMIT_PROJECTED (default)  every job states its length when it
 *     arrives, so the heap use after each future step is known for the live
 *     set: a timeline of projected use per step ahead (up to
 *     KMA_SERVE_MAX_PROJECTED_LEN steps, kept in LDS by the work-group that
 *     admits).  A job is admitted only if adding its growth (1, 2, ..., L
 *     blocks over its L steps) keeps every future step at or below
 *     watermark_pct of the heap's blocks; the first arrival that does not fit
 *     waits, and so do all later ones (first come, first served).  Live jobs
 *     then never find the heap full.
 *   KMA_SERVE_ADMIT_WATERMARK  admit while the superblocks in use are below
 *     watermark_pct of the heap (current use only); a step with stalls
 *     pauses admission until use falls 10 points below the watermark.
 * Under either policy a job whose malloc returns NULL waits a step; after a
 * step with such stalls the most recently admitted jobs are preempted
 * (blocks freed, job requeued ahead of new arrivals, restarting from step
 * 0) until their blocks cover the shortfall (projected admission then
 * rebuilds its timeline from the live set and admits nothing that step).
 * Job j's result is the sum of the 64-bit patterns its blocks carried in the
 * run that completed it, which depends on (j, its length) only. */
#define KMA_SERVE_ADMIT_PROJECTED 0u
#define KMA_SERVE_ADMIT_WATERMARK 1u
#define KMA_SERVE_MAX_PROJECTED_LEN 8192u
typedef struct kma_serve_spec {
	double rate_per_ms;     /* mean arrivals per millisecond */
	double len_mu;          /* log-normal length: exp(mu + sigma z) steps */
	double len_sigma;
	double duration_ms;     /* wall-clock length of the run */
	uint32_t max_len;       /* lengths are capped here (block-table width) */
	uint32_t block_bytes;   /* appended per live job per step */
	uint32_t watermark_pct; /* heap use limit of admission, percent (1..100; projected admission takes up
				   to 400: over 100 it admits more than the heap holds and relies on preemption) */
	uint32_t max_live;      /* job slots (multiple of 256) */
	uint32_t max_admit;     /* admissions per step (0: 1024) */
	uint32_t nwg;           /* work-groups (0: max(64, max_live / 256), co-resident) */
	uint64_t seed;
	uint32_t admit;         /* KMA_SERVE_ADMIT_*; projected needs max_len <= KMA_SERVE_MAX_PROJECTED_LEN */
	uint32_t keep_live;     /* 1: end without freeing the blocks of jobs still live (the heap keeps them, for
				   inspection with kma_snapshot / kma_heap_check; release it with kma_destroy) */
} kma_serve_spec;

typedef struct kma_serve_stats {
	double run_ms;          /* device wall clock from start to the stop decision */
	double kernel_ms;       /* HIP-event time of the whole launch (drain included) */
	uint64_t jobs;          /* jobs in the trace */
	uint64_t truncated;     /* lengths capped at max_len */
	uint64_t steps, admitted, done, drained; /* drained: still live at the stop */
	uint64_t mallocs, frees, stalls, drain_frees, bad;
	uint64_t preempted;     /* admissions ended by preemption (the job is requeued) */
	double ops_per_s;       /* (mallocs + frees) / run_ms */
	double lat_p50_ns, lat_p99_ns, lat_max_ns; /* malloc latency (collective call) */
	double wait_p50_us, wait_p99_us;           /* arrival -> admission, admitted jobs */
	double peak_used_frac;  /* most superblocks out of the pool at a step start */
	uint32_t nwg;
	uint32_t aborted;       /* a barrier watchdog fired */
	uint32_t admit;         /* the admission policy run */
	uint32_t pad;
	uint64_t cap_blocks;    /* projected admission: the block limit (watermark_pct of the heap's blocks) */
	uint64_t rebuilds;      /* projected admission: timeline rebuilds after steps with stalls */
	uint64_t proj_checked;  /* projected admission: steps whose projected use was compared with the blocks
				   held after their growth phase (steps without stalls or preemptions) */
	uint64_t proj_mismatch; /* ... and found different (0 when the projection is exact) */
	double goodput_per_s;   /* completed jobs per second of run_ms */
	double preempt_frac;    /* preempted / admitted */
	double mean_live;       /* live jobs per step, mean over the steps */
	double peak_proj_frac;  /* projected admission: largest projected use of a step / the heap's blocks */
	double phase_us[4];     /* per step, mean: admission work; admission + its grid barrier;
				   growth + barrier; frees + barrier (wall clock of work-group 0) */
	double busy_frac[2];    /* growth, frees: mean fraction of the phase the work-groups spend working */
} kma_serve_stats;

/* Run the serving loop on a native heap.  jobs == NULL: the whole trace;
 * otherwise only the listed job ids, all arriving at the start (with the
 * lengths they have in the whole trace).  results (optional, nresults
 * entries) gets the result of each job of the run (0 if it did not
 * complete): for jobs == NULL indexed by job id, else by position in jobs. */
int kma_serve_run(kma_heap *h, const kma_serve_spec *spec, const uint64_t *jobs, uint64_t njobs,
		  uint64_t *results, uint64_t nresults, kma_serve_stats *out);

/* The serving trace on the host (no device needed): arrival ticks of the
 * device wall clock and lengths of up to `max` jobs; returns the number of
 * jobs in the trace. */
int64_t kma_serve_trace(const kma_serve_spec *spec, uint64_t *arrival, uint32_t *lengths, uint64_t max);

/* Lengths of the listed jobs (host only). */
int kma_serve_lengths(const kma_serve_spec *spec, const uint64_t *jobs, uint64_t n, uint32_t *lengths);

/* ---- graph building: the reference's tb_clTree application ----
 * Every edge {source, sink} makes sure both node ids are in an insert-only
 * tree (tb_clTree.cl:12-38 clGraph_node_ensure), allocates a link and
 * appends it to the source node's adjacency list (tb_clTree.cl:57-67).
 * Nodes (48 B: a tree node and the queue of its links) and links (16 B) come from
 *   KMA_GRAPH_PLAIN       one kma_malloc per lane; a node that lost its
 *                         insert race is freed (clTree_test,
 *                         tb_clTree.cl:41-85);
 *   KMA_GRAPH_CACHE       the same, the lost node kept for the lane's next
 *                         insert (clTree_test_cache, tb_clTree.cl:87-173);
 *                         also runs on the bump allocator (pma != NULL),
 *                         like the reference's "poormans heap" run;
 *   KMA_GRAPH_COLLECTIVE  work-group collective kma_wg_malloc calls for the
 *                         missing source nodes, the missing sink nodes and
 *                         the links (clTree_test_al, tb_clTree.cl:175-273).
 */
#define KMA_GRAPH_PLAIN 0
#define KMA_GRAPH_CACHE 1
#define KMA_GRAPH_COLLECTIVE 2

typedef struct kma_graph_stats {
	double gpu_ms;         /* the build launch */
	uint64_t edges;        /* edges linked */
	uint64_t nodes;        /* distinct node ids in the edge list */
	uint64_t node_mallocs; /* node allocations, those that lost an insert race included */
	uint64_t node_frees;   /* nodes freed after a lost race or left in a lane's cache */
	uint64_t link_mallocs;
	uint64_t nulls;        /* allocations that returned NULL (that lane stops building) */
	uint64_t bad;          /* lost insert races whose winner was not found */
	uint64_t violations;   /* ids missing from the tree, or whose links differ from the edge list */
	uint64_t links_found;  /* links reached from the tree */
	uint32_t nwg;          /* work-groups of 256 used */
	uint32_t pad;
} kma_graph_stats;

/* An edge list of nedges {source, sink} pairs over nnodes distinct ids
 * (pseudo-random, from seed), in edges[2 * nedges].  The reference reads
 * the same pairs from a text file (tb_clTree.c:350-389); none ships. */
int kma_graph_generate(uint32_t nnodes, uint64_t nedges, uint64_t seed, uint32_t *edges);

/* Build the graph of `edges` (n pairs) with one launch of nwg work-groups of
 * 256 (0: one per 256 edges, at most 1,024) on the native heap `heap`, or on
 * the bump allocator `pma` (KMA_GRAPH_CACHE only); then check it against the
 * edge list on the device, and with `teardown` free every node and link
 * (native heap only).  Fills *out; 0 on success. */
struct kma_pma;
int kma_graph_build(kma_heap *heap, struct kma_pma *pma, const uint32_t *edges, uint64_t n, int mode,
		    uint32_t nwg, int teardown, kma_graph_stats *out);

/* ------------------------------------------------------------------ */
/* Heap walker / orphan detector                                        */
/* ------------------------------------------------------------------ */
/* The reference has a commented-out heap walker (kma.cl:528-548) and
 * prints heap dumps from its test bench (tb_kma.c:304-313); this is the
 * walker as a checker.  Every superblock must be in exactly one place:
 *   - free (class 0, count 0, no bit set): in exactly one pool shard chain
 *     or one context spare slot;
 *   - in use: its count equals its set bits, no bit past the class's
 *     slots, its owned bit set iff exactly one context head names it (with
 *     its generation and class), and not in any free list;
 *   - an orphan is a superblock with a class, no live block and no owner:
 *     nobody would ever return it to the pool.
 * Shard counts must equal the chain lengths, a non-empty shard's mask bit
 * must be set, the page bitmap must be a sequence of runs whose first pages
 * record their lengths, and the window summaries of both levels must be
 * exact.  Only valid while no kernel uses the heap.
 *
 * Reference layout (kma.h:71-85): the free queue is walked from its head
 * the way the reference's queue validator walks it
 * (tb_clIndexedQueue.c:131-168): it must end within S nodes (no cycle), at
 * the node the tail names, and every queued superblock must be free (no
 * bit set, free count 0, not a class head).  A class head must be empty or
 * name a superblock of its class (POISON means a refill was cut off).
 * Every superblock out of the queue and in use must count exactly its free
 * slots (slots - set bits), have no bit past its slots and carry a block
 * size of the table; one with no live block that is neither queued nor a
 * class head is an orphan (the reference's commented-out check,
 * kma.cl:528-548).  The last superblock, which the reference's init never
 * queues (kma.cl:42-51), is counted in sb_unqueued, not as a problem. */
typedef struct kma_heap_report {
	uint64_t problems;     /* violations found (0: consistent) */
	uint64_t sb_total;     /* superblocks */
	uint64_t sb_free;      /* free superblocks (pool shards + spare slots) */
	uint64_t sb_in_spares; /* of which in context spare slots */
	uint64_t sb_heads;     /* superblocks attached to a context head */
	uint64_t sb_detached;  /* in use, live blocks, no owner (freed back later) */
	uint64_t orphans;      /* class set, no live block, no owner */
	uint64_t live_blocks;  /* bitmap bits set */
	uint64_t live_bytes;   /* live blocks times their class size */
	uint64_t page_runs;    /* live page runs */
	uint64_t pages_used;
	uint64_t pages_total;
	uint64_t sb_unqueued;  /* reference layout: never queued, never used */
	uint64_t sb_listed;    /* native: detached, partly free, in a partial index */
	uint64_t counters[11]; /* native: the heap's event counters, KMA_STAT_* order;
				  reference: the stats trailer's words 0..7 */
	char first[1024];      /* the first problems, one per line */
} kma_heap_report;

/* Check a metadata image: bytes [0, off_data) of a native heap, or the
 * whole image of a reference-layout heap, as kma_snapshot(h, 0, ...)
 * returns them.  0 with *out filled (problems counted there), -1 if the
 * image is neither heap's. */
int kma_image_check(const void *image, uint64_t bytes, kma_heap_report *out);

/* Snapshot h's metadata and check it (no kernel may be using h). */
int kma_heap_check(kma_heap *h, kma_heap_report *out);

/* ------------------------------------------------------------------ */
/* Deterministic batch mode (reference layout)                          */
/* ------------------------------------------------------------------ */
/* Serve n mallocs as ONE parallel batch whose result is identical, address
 * for address and byte for byte of the heap image, to malloc(sizes[0]), ...,
 * malloc(sizes[n-1]) run one at a time in index order (kma.cl:279-303).  The
 * heap must hold no live block and no attached class superblock (as after
 * kma_create / kma_reset or kma_batch_free_det of the previous batch), and
 * no request may fall in a 1-slot class (serialized, the reference blocks
 * on the second such request).  out[i] = heap offset of request i's block,
 * 0 for NULL.  *ms (optional) = device time of the batch, inputs already
 * resident.  Returns the number of NULLs, or -1. */
int64_t kma_batch_malloc_det(kma_heap *h, const uint32_t *sizes, uint64_t n, uint64_t *out, double *ms);

/* Allocation-free form of kma_batch_malloc_det.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc / hipFree and no allocating library
 * routine, and does not wait: the kernels are enqueued on `stream` (a
 * hipStream_t; NULL = the null stream) and the call returns.
 *   sizes  DEVICE array of n request sizes (input)
 *   out    DEVICE array of n heap offsets (output; 0 = NULL), as
 *          kma_batch_malloc_det returns them
 *   status DEVICE array of 2 u64 (output): [0] the number of NULLs, [1] the
 *          requests refused for falling in a 1-slot class (when nonzero the
 *          heap image is undefined and must be reset, as for the blocking form)
 * Same heap precondition as kma_batch_malloc_det.  Intermediate results
 * live in a static device scratch of kma_det_noalloc_scratch_bytes() bytes
 * per GPU (about 8.9 MiB, held until the process ends; its address on a
 * GPU is resolved when the first reference-layout heap is created there or
 * at the first call).  Every call on one GPU shares that scratch, so calls
 * that overlap on one GPU (other streams, other threads) race; calls in
 * order on one stream do not.  -1 with kma_last_error() for n == 0,
 * n > KMA_DET_NOALLOC_MAX, a native-layout heap or a failed launch. */
#define KMA_DET_NOALLOC_MAX 524288u
int kma_batch_malloc_det_noalloc(kma_heap *h, const uint32_t *sizes, uint64_t n, uint64_t *out,
				 uint64_t *status, void *stream);
uint64_t kma_det_noalloc_scratch_bytes(void);

/* Free the blocks of the preceding kma_batch_malloc_det batch (offs = its
 * out array, n its length) with the result of free(offs[0]), ...,
 * free(offs[n-1]) run one at a time in index order (kma.cl:305-373). */
int kma_batch_free_det(kma_heap *h, const uint64_t *offs, uint64_t n, double *ms);

#ifdef __cplusplus
}
#endif
#endif
