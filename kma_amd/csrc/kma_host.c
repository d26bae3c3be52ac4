/*
 * kma_host.c - host side of libkma.so, in C over the HIP runtime C API.
 *
 * Replaces the reference's OpenCL host constructors (kma.c:31-127): heap
 * buffer allocation, header write and the heap-init launch, plus the
 * workload runners the tests and benchmarks call.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kma.h"
#include "kma_launch.h"
#include "kma_layout.h"
#include "kma_work.h"

#include "kma_internal.h"

static __thread char g_err[256];

void kma_i_set_err(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
}

const char *kma_last_error(void)
{
	return g_err;
}

int kma_device_ok(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
		set_err("no HIP device");
		return -1;
	}
	return 0;
}

/* ---------------- reference layout ---------------- */
static kma_heap *create_reference(const kma_config *cfg)
{
	uint32_t log2 = cfg->sb_log2;
	if (log2 < 8 || log2 > 16) {
		set_err("reference layout: sb_log2 %u outside 8..16", log2);
		return NULL;
	}
	if (cfg->sblocks < 2 || cfg->sblocks > 65536) {
		/* 16-bit effective queue index, clIndexedQueue.cl:89 */
		set_err("reference layout: sblocks %llu outside 2..65536",
			(unsigned long long)cfg->sblocks);
		return NULL;
	}
	kma_heap *h = (kma_heap *)calloc(1, sizeof(*h));
	if (!h)
		return NULL;
	h->cfg = *cfg;
	uint32_t hdr = 32u + 8u * (log2 - 1u);
	h->bytes = (cfg->sblocks << log2) + hdr;
	/* the image, then the stats trailer the device code finds after it */
	if (hipMalloc((void **)&h->dev, REF_TRAILER_OFF(h->bytes) + REF_TRAILER_BYTES) != hipSuccess) {
		set_err("KMA: Could not allocate heap on-device (%llu bytes)", (unsigned long long)h->bytes);
		free(h);
		return NULL;
	}
	if (kma_reset(h) != 0) {
		hipFree(h->dev);
		free(h);
		return NULL;
	}
	/* resolve this GPU's scratch address for the allocation-free batch
	 * entry point now, so that its calls need not */
	int rc = kmad_det_scratch_setup(NULL);
	if (rc != 0) {
		set_err("scratch address of the allocation-free batch: %s", hipGetErrorString((hipError_t)(-rc)));
		hipFree(h->dev);
		free(h);
		return NULL;
	}
	return h;
}

/* ---------------- native layout ---------------- */
static int native_layout(kma_heap *h, const kma_config *in)
{
	char err[200];
	if (kma_native_layout(in, &h->cfg, &h->nh, err, sizeof(err)) != 0) {
		set_err("%s", err);
		return -1;
	}
	return 0;
}

static kma_heap *create_native(const kma_config *in)
{
	kma_heap *h = (kma_heap *)calloc(1, sizeof(*h));
	if (!h)
		return NULL;
	if (native_layout(h, in) != 0) {
		free(h);
		return NULL;
	}
	nat_hdr *nh = &h->nh;
	uint64_t total = kma_native_alloc_bytes(nh);
	if (hipMalloc((void **)&h->dev, total) != hipSuccess) {
		set_err("KMA: Could not allocate heap on-device (%llu bytes)", (unsigned long long)total);
		free(h);
		return NULL;
	}
	h->bytes = total;
	kma_native_place(nh, (uint64_t)(uintptr_t)h->dev, total);
	if (kma_reset(h) != 0) {
		hipFree(h->dev);
		free(h);
		return NULL;
	}
	return h;
}

int kma_reset(kma_heap *h)
{
	if (!h)
		return -1;
	if (h->cfg.layout == KMA_LAYOUT_REFERENCE) {
		/* The reference does not clear its buffer (kma.c:38); the heap
		 * image this library defines starts from zero (the stats trailer
		 * and its deadline too). */
		HIPCHK(hipMemset(h->dev, 0, REF_TRAILER_OFF(h->bytes) + REF_TRAILER_BYTES), -1);
		LCHK(kmad_ref_init(h->dev, h->cfg.sb_log2, (uint32_t)h->cfg.sblocks), -1);
		return 0;
	}
	nat_hdr *nh = &h->nh;
	HIPCHK(hipMemset(h->dev, 0, nh->off_data), -1);
	HIPCHK(hipMemcpy(h->dev, nh, sizeof(*nh), hipMemcpyHostToDevice), -1);
	LCHK(kmad_nat_init(h->dev, nh->nsb, nh->npool, (uint64_t)nh->nctx * NAT_HEADS_PER_CTX,
			   (nh->npages + 63) / 64), -1);
	return 0;
}

kma_heap *kma_create_ex(const kma_config *cfg)
{
	if (!cfg) {
		set_err("NULL config");
		return NULL;
	}
	if (kma_device_ok() != 0)
		return NULL;
	if (cfg->layout == KMA_LAYOUT_REFERENCE)
		return create_reference(cfg);
	if (cfg->layout == KMA_LAYOUT_NATIVE)
		return create_native(cfg);
	set_err("unknown layout %u", cfg->layout);
	return NULL;
}

kma_heap *kma_create(unsigned int sblocks)
{
	/* kma.c:89-90 */
	if (sblocks == 0) {
		set_err("sblocks == 0");
		return NULL;
	}
	kma_config c;
	memset(&c, 0, sizeof(c));
	c.layout = KMA_LAYOUT_REFERENCE;
	c.sb_log2 = 12; /* KMA_SB_SIZE, kma.h:27-28 */
	c.sblocks = sblocks;
	return kma_create_ex(&c);
}

void kma_destroy(kma_heap *h)
{
	if (!h)
		return;
	if (h->dev)
		hipFree(h->dev);
	free(h);
}

void *kma_device_heap(const kma_heap *h)
{
	return h ? h->dev : NULL;
}

uint64_t kma_heap_bytes(const kma_heap *h)
{
	return h ? h->bytes : 0;
}

int kma_get_config(const kma_heap *h, kma_config *out)
{
	if (!h || !out)
		return -1;
	*out = h->cfg;
	return 0;
}

int kma_snapshot(const kma_heap *h, uint64_t offset, void *dst, uint64_t bytes)
{
	if (!h || !dst || offset + bytes > h->bytes) {
		set_err("snapshot out of range");
		return -1;
	}
	HIPCHK(hipMemcpy(dst, h->dev + offset, bytes, hipMemcpyDeviceToHost), -1);
	return 0;
}

/* ---------------- workloads ---------------- */
int kma_trace_serial(kma_heap *h, const int64_t *ops, uint64_t n, uint64_t *out, const uint32_t *pat)
{
	if (!h || (!ops && n) || (!out && n)) {
		set_err("bad arguments");
		return -1;
	}
	if (n == 0)
		return 0;
	int64_t *d_ops = NULL;
	uint64_t *d_out = NULL;
	uint32_t *d_pat = NULL;
	int rc = -1;
	if (hipMalloc((void **)&d_ops, n * 8) != hipSuccess || hipMalloc((void **)&d_out, n * 8) != hipSuccess ||
	    (pat && hipMalloc((void **)&d_pat, n * 4) != hipSuccess)) {
		set_err("trace buffers: hipMalloc failed");
		goto out;
	}
	for (uint64_t i = 0; i < n; i++) {
		if (ops[i] < 0 && (uint64_t)(-ops[i] - 1) >= i) {
			set_err("op %llu frees a later op", (unsigned long long)i);
			goto out;
		}
	}
	if (hipMemcpy(d_ops, ops, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemset(d_out, 0, n * 8) != hipSuccess ||
	    (pat && hipMemcpy(d_pat, pat, n * 4, hipMemcpyHostToDevice) != hipSuccess)) {
		set_err("trace upload failed");
		goto out;
	}
	rc = kmad_serial(h->cfg.layout == KMA_LAYOUT_REFERENCE ? 0 : 1, h->dev, d_ops, n, d_out, d_pat);
	if (rc != 0) {
		set_err("serial trace kernel failed: %d", rc);
		rc = -1;
		goto out;
	}
	if (hipMemcpy(out, d_out, n * 8, hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("trace download failed");
		rc = -1;
	}
out:
	if (d_ops)
		hipFree(d_ops);
	if (d_out)
		hipFree(d_out);
	if (d_pat)
		hipFree(d_pat);
	return rc;
}

int kma_trim(kma_heap *h)
{
	if (!h || h->cfg.layout != KMA_LAYOUT_NATIVE)
		return 0;
	LCHK(kmad_trim(h->dev, h->nh.nctx, h->nh.nclasses, (uint32_t)((h->nh.nsb + 63) / 64)), -1);
	return 0;
}

int kma_i_read_native_stats(kma_heap *h, uint64_t *st)
{
	if (h->cfg.layout != KMA_LAYOUT_NATIVE) {
		memset(st, 0, NAT_NSTATS * 8);
		return 0;
	}
	HIPCHK(hipMemcpy(st, h->dev + h->nh.off_stats, NAT_NSTATS * 8, hipMemcpyDeviceToHost), -1);
	return 0;
}

int kma_i_read_ref_trailer(kma_heap *h, uint64_t *tw)
{
	if (h->cfg.layout != KMA_LAYOUT_REFERENCE) {
		memset(tw, 0, REF_T_WORDS * 8);
		return 0;
	}
	HIPCHK(hipMemcpy(tw, h->dev + REF_TRAILER_OFF(h->bytes), REF_T_WORDS * 8, hipMemcpyDeviceToHost), -1);
	return 0;
}

static int cmp_u32(const void *a, const void *b)
{
	uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
	return x < y ? -1 : x > y;
}

/* nanoseconds per tick of the device's constant wall clock */
double kma_i_wall_clock_ns(void)
{
	int dev = 0, khz = 0;
	if (hipGetDevice(&dev) != hipSuccess ||
	    hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
		return 10.0; /* 100 MHz */
	return 1e6 / (double)khz;
}

static int launch_form(kma_heap *h, const kma_workload *wl, uint32_t form, uint32_t round0, uint32_t nrounds,
		       uint32_t variant, unsigned long long *d_stats, uint32_t *d_lat, hipStream_t s)
{
	int layout = h->cfg.layout == KMA_LAYOUT_REFERENCE ? 0 : 1;
	if (form == KMA_FORM_BATCHED)
		return kmad_batched(h->dev, round0, nrounds, (int)wl->mode, wl->fixed, wl->seed, wl->nwi, variant,
				    d_stats, d_lat, s);
	return kmad_rounds(layout, h->dev, round0, nrounds, (int)wl->mode, wl->fixed, wl->seed, wl->nwi, wl->block,
			   variant, d_stats, d_lat, s);
}

int kma_run_workload(kma_heap *h, const kma_workload *wl, kma_run_stats *out)
{
	if (!h || !wl || !out) {
		set_err("bad arguments");
		return -1;
	}
	memset(out, 0, sizeof(*out));
	if (wl->mode > KMA_WL_HIGHVAR || wl->form > KMA_FORM_FUSED || wl->nwi == 0 || wl->block == 0 ||
	    (wl->block & 63) || wl->block > 256 || wl->rounds == 0) {
		set_err("bad workload");
		return -1;
	}
	if (wl->form == KMA_FORM_BATCHED && h->cfg.layout != KMA_LAYOUT_NATIVE) {
		set_err("the batched form needs a native-layout heap");
		return -1;
	}
	unsigned long long *d_stats = NULL;
	uint32_t *d_lat = NULL, *h_lat = NULL;
	hipStream_t s = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	hipGraph_t graph = NULL;
	hipGraphExec_t gexec = NULL;
	int rc = -1;
	uint64_t st0[NAT_NSTATS], st1[NAT_NSTATS], tw0[REF_T_WORDS], tw1[REF_T_WORDS];
	int ref = h->cfg.layout == KMA_LAYOUT_REFERENCE;
	uint64_t budget = 0;
	if (ref && wl->deadline_ms)
		budget = (uint64_t)((double)wl->deadline_ms * 1e6 / kma_i_wall_clock_ns());
	unsigned long long hs[8];
	uint64_t nlat = (wl->variant & KMA_VAR_LATENCY) ? wl->nwi * wl->rounds : 0;
	uint32_t untimed = wl->variant & ~(KMA_VAR_LATENCY | KMA_VAR_PHASES);
	if (hipMalloc((void **)&d_stats, sizeof(hs)) != hipSuccess || hipStreamCreate(&s) != hipSuccess ||
	    hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
	    (nlat && hipMalloc((void **)&d_lat, nlat * 4) != hipSuccess)) {
		set_err("workload setup failed");
		goto done;
	}
	if (hipMemset(d_stats, 0, sizeof(hs)) != hipSuccess)
		goto done;
	/* warm-up rounds, untimed, not counted */
	if (wl->warmup) {
		rc = launch_form(h, wl, wl->form == KMA_FORM_PLAIN ? KMA_FORM_FUSED : wl->form, wl->round0,
				 wl->warmup, untimed, d_stats, NULL, s);
		if (rc) {
			set_err("warm-up launch failed: %d", rc);
			goto done;
		}
	}
	if (hipStreamSynchronize(s) != hipSuccess || hipMemset(d_stats, 0, sizeof(hs)) != hipSuccess ||
	    kma_i_read_native_stats(h, st0) != 0 || kma_i_read_ref_trailer(h, tw0) != 0) {
		set_err("warm-up failed");
		rc = -1;
		goto done;
	}
	uint32_t r0 = wl->round0 + wl->warmup;
	if (wl->form == KMA_FORM_BATCHED) {
		/* one persistent launch for all rounds, captured in a graph */
		if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
			set_err("graph capture failed");
			rc = -1;
			goto done;
		}
		rc = launch_form(h, wl, KMA_FORM_BATCHED, r0, wl->rounds, wl->variant, d_stats, d_lat, s);
		if (hipStreamEndCapture(s, &graph) != hipSuccess || rc != 0 ||
		    hipGraphInstantiate(&gexec, graph, NULL, NULL, 0) != hipSuccess) {
			set_err("graph build failed: %d", rc);
			rc = -1;
			goto done;
		}
	}
	/* reference layout: the deadline runs from here over every launch of
	 * the timed region (every wait gives up once it has passed) */
	if (budget && (rc = kmad_ref_arm(h->dev, budget, s)) != 0) {
		set_err("deadline: launch failed: %d", rc);
		goto done;
	}
	hipEventRecord(e0, s);
	if (wl->form == KMA_FORM_PLAIN) {
		for (uint32_t r = 0; r < wl->rounds; r++) {
			rc = kmad_rounds(h->cfg.layout == KMA_LAYOUT_REFERENCE ? 0 : 1, h->dev, r0 + r, 1, (int)wl->mode,
					 wl->fixed, wl->seed, wl->nwi, wl->block, wl->variant, d_stats,
					 d_lat ? d_lat + (uint64_t)r * wl->nwi : NULL, s);
			if (rc) {
				set_err("launch failed: %d", rc);
				goto done;
			}
		}
		out->launches = wl->rounds;
	} else if (wl->form == KMA_FORM_BATCHED) {
		if (hipGraphLaunch(gexec, s) != hipSuccess) {
			set_err("graph launch failed");
			rc = -1;
			goto done;
		}
		out->launches = 1;
	} else {
		rc = launch_form(h, wl, KMA_FORM_FUSED, r0, wl->rounds, wl->variant, d_stats, d_lat, s);
		if (rc) {
			set_err("launch failed: %d", rc);
			goto done;
		}
		out->launches = 1;
	}
	hipEventRecord(e1, s);
	if (budget && (rc = kmad_ref_arm(h->dev, 0, s)) != 0) {
		set_err("deadline: launch failed: %d", rc);
		goto done;
	}
	if (hipEventSynchronize(e1) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
		set_err("workload kernels failed: %s", hipGetErrorString(hipGetLastError()));
		rc = -1;
		goto done;
	}
	float ms = 0;
	hipEventElapsedTime(&ms, e0, e1);
	out->gpu_ms = ms;
	out->kernel_ms_avg = out->launches ? ms / (double)out->launches : 0;
	if (hipMemcpy(hs, d_stats, sizeof(hs), hipMemcpyDeviceToHost) != hipSuccess || kma_i_read_native_stats(h, st1) != 0 ||
	    kma_i_read_ref_trailer(h, tw1) != 0) {
		set_err("stats download failed");
		rc = -1;
		goto done;
	}
	out->ops = hs[0];
	out->nulls = hs[1];
	out->bad = hs[2];
	out->mallocs = hs[3];
	out->double_frees = st1[NAT_STAT_DOUBLE_FREE] - st0[NAT_STAT_DOUBLE_FREE];
	out->bad_frees = st1[NAT_STAT_BAD_FREE] - st0[NAT_STAT_BAD_FREE];
	out->pool_empty = st1[2] - st0[2];
	for (int k = 0; k < 4; k++)
		out->waits[k] = st1[3 + k] - st0[3 + k];
	if (ref) {
		/* trailer words 1..5: reserve watchdog, deadline, claim, dequeue,
		 * enqueue */
		out->waits[0] = tw1[1] - tw0[1];
		out->waits[1] = tw1[3] - tw0[3];
		out->waits[2] = tw1[4] - tw0[4];
		out->waits[3] = tw1[5] - tw0[5];
		out->deadline_hits = tw1[2] - tw0[2];
	}
	out->remote_frees = st1[7] - st0[7];
	out->migrated = st1[8] - st0[8];
	out->listed = st1[9] - st0[9];
	out->reattached = st1[10] - st0[10];
	double tick = kma_i_wall_clock_ns();
	if (hs[7])
		for (int k = 0; k < 3; k++)
			out->phase_ns[k] = tick * (double)hs[4 + k] / (double)hs[7];
	if (nlat) {
		h_lat = (uint32_t *)malloc(nlat * 4);
		if (!h_lat || hipMemcpy(h_lat, d_lat, nlat * 4, hipMemcpyDeviceToHost) != hipSuccess) {
			set_err("latency download failed");
			rc = -1;
			goto done;
		}
		qsort(h_lat, nlat, 4, cmp_u32);
		out->lat_p50_ns = tick * h_lat[nlat / 2];
		out->lat_p99_ns = tick * h_lat[(nlat * 99) / 100];
		out->lat_max_ns = tick * h_lat[nlat - 1];
	}
	rc = 0;
done:
	free(h_lat);
	if (gexec)
		hipGraphExecDestroy(gexec);
	if (graph)
		hipGraphDestroy(graph);
	if (e0)
		hipEventDestroy(e0);
	if (e1)
		hipEventDestroy(e1);
	if (s)
		hipStreamDestroy(s);
	if (d_lat)
		hipFree(d_lat);
	if (d_stats)
		hipFree(d_stats);
	return rc;
}

int kma_atomic_peak(uint64_t bytes, uint64_t nthreads, uint32_t iters, int pattern, uint32_t warmup,
		    uint32_t reps, double *ops_per_s, double *ms_out)
{
	if (bytes < 4096 || nthreads == 0 || iters == 0 || reps == 0 || pattern < 0 || pattern > 3) {
		set_err("bad arguments");
		return -1;
	}
	if (kma_device_ok() != 0)
		return -1;
	void *buf = NULL;
	hipStream_t s = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	int rc = -1;
	if (hipMalloc(&buf, bytes) != hipSuccess || hipMemset(buf, 0, bytes) != hipSuccess ||
	    hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
	    hipEventCreate(&e1) != hipSuccess) {
		set_err("atomic peak setup failed");
		goto done;
	}
	for (uint32_t i = 0; i < warmup; i++)
		if ((rc = kmad_atomic_peak(buf, bytes / 4, nthreads, iters, pattern, s)) != 0)
			goto done;
	hipEventRecord(e0, s);
	for (uint32_t i = 0; i < reps; i++)
		if ((rc = kmad_atomic_peak(buf, bytes / 4, nthreads, iters, pattern, s)) != 0)
			goto done;
	hipEventRecord(e1, s);
	if (hipEventSynchronize(e1) != hipSuccess) {
		set_err("atomic peak kernel failed");
		rc = -1;
		goto done;
	}
	float ms = 0;
	hipEventElapsedTime(&ms, e0, e1);
	if (ops_per_s)
		*ops_per_s = (double)nthreads * iters * reps / (ms * 1e-3);
	if (ms_out)
		*ms_out = ms / reps;
	rc = 0;
done:
	if (e0)
		hipEventDestroy(e0);
	if (e1)
		hipEventDestroy(e1);
	if (s)
		hipStreamDestroy(s);
	if (buf)
		hipFree(buf);
	return rc;
}

int kma_wave_free(kma_heap *h, const uint64_t *offs, uint32_t n)
{
	if (!h || h->cfg.layout != KMA_LAYOUT_NATIVE || !offs || n == 0 || n > 64) {
		set_err("bad arguments (native heap, 1..64 offsets)");
		return -1;
	}
	for (uint32_t i = 0; i < n; i++)
		if (offs[i] && (offs[i] < h->nh.off_data || offs[i] >= h->bytes)) {
			set_err("offset %u outside the heap's data", i);
			return -1;
		}
	uint64_t *d = NULL;
	if (hipMalloc((void **)&d, n * 8ull) != hipSuccess) {
		set_err("hipMalloc failed");
		return -1;
	}
	int rc = -1;
	if (hipMemcpy(d, offs, n * 8ull, hipMemcpyHostToDevice) != hipSuccess)
		set_err("upload failed");
	else if (kmad_wave_free(h->dev, d, n) != 0)
		set_err("free kernel failed");
	else
		rc = 0;
	hipFree(d);
	return rc;
}

int kma_check_wave_claim(kma_heap *h, uint32_t n, const uint32_t *cls, const uint32_t *base, const uint32_t *got,
			 const uint32_t *bm_in, uint64_t *res_out, uint32_t *bm_out, uint32_t *done_out)
{
	if (!h || h->cfg.layout != KMA_LAYOUT_NATIVE || n == 0 || n > h->nh.nsb) {
		set_err("bad arguments");
		return -1;
	}
	uint32_t W = h->nh.bm_words;
	for (uint32_t i = 0; i < n; i++) {
		if (cls[i] >= h->nh.nclasses || got[i] > 64) {
			set_err("case %u: class or count out of range", i);
			return -1;
		}
	}
	void *d[7] = {0};
	size_t sz[7] = {n * 4ull, n * 4ull, n * 4ull, (size_t)n * W * 4, n * 64ull * 8, (size_t)n * W * 4, n * 4ull};
	int rc = -1;
	for (int k = 0; k < 7; k++)
		if (hipMalloc(&d[k], sz[k]) != hipSuccess) {
			set_err("hipMalloc failed");
			goto out;
		}
	if (hipMemcpy(d[0], cls, sz[0], hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(d[1], base, sz[1], hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(d[2], got, sz[2], hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(d[3], bm_in, sz[3], hipMemcpyHostToDevice) != hipSuccess) {
		set_err("upload failed");
		goto out;
	}
	rc = kmad_check_wave_claim(h->dev, n, d[0], d[1], d[2], d[3], d[4], d[5], d[6]);
	if (rc) {
		set_err("check kernel failed: %d", rc);
		rc = -1;
		goto out;
	}
	if (hipMemcpy(res_out, d[4], sz[4], hipMemcpyDeviceToHost) != hipSuccess ||
	    hipMemcpy(bm_out, d[5], sz[5], hipMemcpyDeviceToHost) != hipSuccess ||
	    hipMemcpy(done_out, d[6], sz[6], hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("download failed");
		rc = -1;
	}
out:
	for (int k = 0; k < 7; k++)
		if (d[k])
			hipFree(d[k]);
	return rc;
}

int kma_copy_probe(uint64_t bytes, uint32_t warmup, uint32_t reps, double *gbps, double *ms_out)
{
	if (bytes < 4096 || (bytes & 15) || reps == 0) {
		set_err("bad arguments");
		return -1;
	}
	if (kma_device_ok() != 0)
		return -1;
	void *a = NULL, *b = NULL;
	hipStream_t s = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	int rc = -1;
	if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
	    hipMemset(a, 1, bytes) != hipSuccess || hipStreamCreate(&s) != hipSuccess ||
	    hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		set_err("copy probe setup failed");
		goto done;
	}
	for (uint32_t i = 0; i < warmup; i++)
		if ((rc = kmad_copy(a, b, bytes, s)) != 0)
			goto done;
	hipEventRecord(e0, s);
	for (uint32_t i = 0; i < reps; i++)
		if ((rc = kmad_copy(a, b, bytes, s)) != 0)
			goto done;
	hipEventRecord(e1, s);
	if (hipEventSynchronize(e1) != hipSuccess) {
		set_err("copy kernel failed");
		rc = -1;
		goto done;
	}
	float ms = 0;
	hipEventElapsedTime(&ms, e0, e1);
	if (gbps)
		*gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
	if (ms_out)
		*ms_out = ms / reps;
	rc = 0;
done:
	if (e0)
		hipEventDestroy(e0);
	if (e1)
		hipEventDestroy(e1);
	if (s)
		hipStreamDestroy(s);
	if (a)
		hipFree(a);
	if (b)
		hipFree(b);
	return rc;
}
