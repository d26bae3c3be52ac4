/*
 * kma_launch.h - internal C prototypes between the host C layer (kma_host.c)
 * and the HIP kernels (kma_kernels.hip).  Plain C types only.
 * Every function returns 0 or a negative hipError_t value.
 */
#ifndef KMA_LAUNCH_H
#define KMA_LAUNCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference layout: parallel clheap_init producing the serial init's image. */
int kmad_ref_init(void *heap, uint32_t log2, uint32_t sblocks);

/* Reference layout: set the stats trailer's deadline to the device wall
 * clock + budget_ticks when the kernel runs on `stream` (0: no deadline). */
int kmad_ref_arm(void *heap, uint64_t budget_ticks, void *stream);

/* Native layout: metadata init (header already copied to device). */
int kmad_nat_init(void *heap, uint64_t nsb, uint32_t npool, uint64_t nheads,
		  uint64_t npbwords);

/* Serialized trace on one work-item. layout selects the device API. */
int kmad_serial(int layout, void *heap, const int64_t *d_ops, uint64_t n,
		uint64_t *d_out, const uint32_t *d_pat);

/* Workload kernels: see kma_work.h for the parameters. */
int kmad_rounds(int layout, void *heap, uint32_t round0, uint32_t nrounds,
		int mode, uint32_t fixed, uint64_t seed, uint64_t nwi,
		uint32_t block, uint32_t variant, unsigned long long *d_stats,
		uint32_t *d_lat, void *stream);

int kmad_batched(void *heap, uint32_t round0, uint32_t nrounds, int mode,
		 uint32_t fixed, uint64_t seed, uint64_t nwi, uint32_t variant,
		 unsigned long long *d_stats, uint32_t *d_lat, void *stream);

int kmad_trim(void *heap, uint32_t nctx, uint32_t nclasses, uint32_t pidx_w0);

int kmad_wave_free(void *heap, const uint64_t *offs, uint32_t n);
int kmad_check_wave_claim(void *heap, uint32_t ncases, const uint32_t *cls, const uint32_t *base,
			  const uint32_t *got, const uint32_t *bm_in, uint64_t *res_out,
			  uint32_t *bm_out, uint32_t *done_out);

int kmad_atomic_peak(void *buf, uint64_t nwords, uint64_t nthreads, uint32_t iters,
		     int pattern, void *stream);
int kmad_copy(const void *src, void *dst, uint64_t bytes, void *stream);

/* Batched request streams (kma_stream.hip).  ctl is a device array of
 * KMA_CTL_WORDS u64: epoch, chunk counter and counters. */
#define KMA_STREAM_BLOCK 256
#define KMA_CTL_EPOCH 0
#define KMA_CTL_NEXT 16
#define KMA_CTL_STATS 32
#define KMA_CTL_WORDS 64

int kmad_stream_begin(uint64_t *d_ctl, void *stream);
int kmad_stream_plain(void *heap, const uint32_t *d_req, uint64_t *d_res, uint64_t lo, uint64_t n,
		      uint64_t *d_ctl, uint64_t *d_ts, uint32_t variant, void *stream);
int kmad_stream_persistent(void *heap, const uint32_t *d_req, uint64_t *d_res, uint64_t nreq,
			   uint32_t nwg, uint32_t per_thread, uint64_t *d_ctl, uint64_t *d_ts,
			   uint32_t variant, void *stream);

/* Containers (kma_containers.hip). */
int kmad_pma_init(void *h, uint64_t bytes, uint64_t align);
int kmad_pma_run(void *h, const uint32_t *d_sizes, uint64_t n, int serial, uint64_t *d_out);
int kmad_queue_init(void *q, uint64_t cap);
int kmad_lqueue_init(void *q);
int kmad_lqueue_run(void *q, uint32_t mode, char *d_items, uint64_t n, uint32_t rounds, int64_t *d_order,
		    unsigned long long *d_lost);
int kmad_queue_run(void *q, uint32_t mode, char *d_items, uint64_t n, uint32_t rounds, int64_t *d_order);
int kmad_list_grow(void *l, const uint32_t *d_objs, uint64_t n, uint32_t block, int local, uint64_t *d_out);
int kmad_list_get(void *l, const uint64_t *d_idx, uint64_t n, uint64_t *d_out);
int kmad_list_clear(void *l, uint64_t n);
int kmad_graph_build(void *heap, void *pma, void *tree, const uint32_t *d_edges, uint64_t n, int mode,
		     uint32_t nwg, unsigned long long *d_stats, void *stream);
int kmad_graph_check(void *tree, const uint32_t *d_keys, const uint32_t *d_deg, const uint64_t *d_ssum,
		     uint64_t nk, uint64_t *d_nodes, unsigned long long *d_stats, void *stream);
int kmad_graph_free(void *heap, const uint64_t *d_nodes, const uint32_t *d_deg, uint64_t nk, void *stream);
int kmad_tree_run(void *heap, void *tree, const uint32_t *d_keys, uint64_t n, uint32_t *d_added,
		  const uint32_t *d_query, uint64_t nq, uint32_t *d_found, void *d_stack,
		  uint32_t *d_inorder, uint64_t *d_res);

/* Serving loop (kma_serve.hip).  ctl: KMA_SV_CTL_WORDS u64. */
#define KMA_SV_CTL_WORDS 256
#define KMA_SV_STATS 160 /* ctl index of the counters */
#define KMA_SV_NHIST 65536
typedef struct kma_serve_args {
	void *heap;               /* kma_dheap */
	const uint64_t *arrival;  /* per job: arrival tick after the loop's start */
	const uint32_t *length;   /* per job: steps */
	const uint64_t *job_id;   /* per job: the id its patterns use (NULL: its index) */
	uint64_t njobs;
	uint32_t max_live;        /* job slots, multiple of 256 */
	uint32_t max_len;         /* block-table width */
	uint32_t block_bytes;
	uint32_t watermark_pct;
	uint32_t max_admit;       /* admissions per step */
	uint32_t admit;           /* KMA_SERVE_ADMIT_* */
	uint64_t duration_ticks;
	uint64_t cap_blocks;      /* projected admission: block limit (< 2^31) */
	uint32_t keep_live;       /* skip the drain */
	uint32_t no_partial;
	int64_t *slot_job;        /* -1: free */
	uint32_t *slot_steps;
	uint64_t *table;          /* [max_live][max_len] block addresses */
	uint32_t *free_slots;     /* stack, max_live entries */
	uint64_t *ctl;
	uint64_t *done;           /* [max_live][4]: slot, job, first, count */
	uint64_t *result;         /* per job: sum of its blocks' patterns */
	uint64_t *job_admit;      /* per job: admission tick + 1 (0: never) */
	uint64_t *job_done;       /* per job: completion tick + 1 (0: never) */
	uint64_t *lat_hist;       /* KMA_SV_NHIST buckets of 1 wall-clock tick */
	uint32_t *slot_seq;       /* per slot: admission sequence number */
	uint64_t *lstack;         /* admission-ordered (seq << 32 | slot) entries */
	uint64_t lstack_cap;      /* >= 2 * max_live */
	uint64_t *requeue;        /* ring of preempted job ids */
	uint64_t rq_cap;
} kma_serve_args;

/* Largest co-resident grid for the serving kernel (256-thread groups). */
int kmad_serve_max_groups(uint32_t *out);
int kmad_serve_launch(const kma_serve_args *a, uint32_t nwg, void *stream);

/* Deterministic batch mode, reference layout (kma_detbatch.hip). */
int kmad_det_scan(uint32_t *d_tab, uint64_t nrows, uint32_t ncol, uint32_t *d_tot);
/* Buffers of one deterministic malloc batch of n requests (nb = ceil(n/256)):
 * cls[n], rank[n], cnt[nb*16], tot[16], evcnt[nb], small[64] ([0] refill
 * events, [1] successful pops, [2] refused requests, [16..32) class offsets),
 * table[n+16], evc[n], evq[n], popped[min(n, 65536)]. */
typedef struct kmad_det_bufs {
	int8_t *cls;
	uint32_t *rank, *cnt, *tot, *evcnt, *small, *table, *evc, *evq, *popped;
} kmad_det_bufs;
/* Every kernel of one malloc batch, in order, on `stream`. */
int kmad_det_malloc(void *heap, const uint32_t *d_sizes, uint64_t n, const kmad_det_bufs *b, uint64_t *d_out,
		    unsigned long long *d_nulls, void *stream);
/* Allocation-free form: the buffers are carved from the library's static
 * device scratch (one per GPU).  kmad_det_scratch_setup resolves that GPU's
 * scratch address (current device) once; *out (optional) receives it. */
int kmad_det_scratch_setup(void **out);
uint64_t kmad_det_scratch_bytes(void);
int kmad_det_malloc_noalloc(void *heap, const uint32_t *d_sizes, uint64_t n, uint64_t *d_out,
			    unsigned long long *d_status, void *stream);
int kmad_detf_min(void *heap, const uint64_t *d_offs, uint64_t n, uint32_t *d_minblk);
int kmad_detf_words(void *heap, const uint64_t *d_offs, uint64_t n, const uint32_t *d_minblk,
		    uint32_t *d_evcnt, uint32_t *d_last_sb);
int kmad_detf_order(void *heap, const uint64_t *d_offs, uint64_t n, const uint32_t *d_minblk,
		    const uint32_t *d_evoff, const uint32_t *d_last_sb, uint32_t *d_order);
int kmad_detf_enqueue(void *heap, const uint32_t *d_order, const uint32_t *d_ne);

#ifdef __cplusplus
}
#endif
#endif
