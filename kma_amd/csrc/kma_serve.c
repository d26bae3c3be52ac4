/*
 * kma_serve.c - host side of the serving loop (kma_serve_run): job trace
 * generation, device buffers, the cooperative launch and the statistics.
 * The loop itself is in kma_serve.hip.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "kma.h"
#include "kma_internal.h"
#include "kma_launch.h"
#include "kma_work.h"

#define SV_C_NEXT 32
#define SV_C_ABORT 112

static uint64_t smix(uint64_t seed, uint64_t a, uint64_t b)
{
	uint64_t z = seed * 0x9E3779B97F4A7C15ull + a * 0xBF58476D1CE4E5B9ull + b * 0x94D049BB133111EBull;
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

/* uniform in (0, 1] */
static double unit(uint64_t x)
{
	return ((double)(x >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

/* Job j's length: exp(mu + sigma z), z standard normal by Box-Muller from
 * two uniforms of (seed, j); rounded, at least 1, capped at max_len. */
static uint32_t job_length(const kma_serve_spec *sp, uint64_t j, int *capped)
{
	double u1 = unit(smix(sp->seed, j, 2)), u2 = unit(smix(sp->seed, j, 3));
	double z = sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
	double L = floor(exp(sp->len_mu + sp->len_sigma * z) + 0.5);
	*capped = L > sp->max_len;
	if (L < 1)
		L = 1;
	if (L > sp->max_len)
		L = sp->max_len;
	return (uint32_t)L;
}

/* Arrival ticks (exponential gaps of mean ticks_per_ms / rate from
 * uniforms of (seed, j), until the duration) and lengths of the trace, or of
 * the listed jobs (arriving at 0). */
static int gen_trace(const kma_serve_spec *sp, const uint64_t *jobs, uint64_t njobs_sel, double ticks_per_ms,
		     uint64_t **arr_out, uint32_t **len_out, uint64_t *n_out, uint64_t *truncated)
{
	uint64_t cap = jobs ? (njobs_sel ? njobs_sel : 1) : 1024, n = 0;
	uint64_t *arr = (uint64_t *)malloc(cap * 8);
	uint32_t *len = (uint32_t *)malloc(cap * 4);
	if (!arr || !len)
		goto oom;
	if (jobs) {
		for (uint64_t i = 0; i < njobs_sel; i++) {
			int c;
			arr[i] = 0;
			len[i] = job_length(sp, jobs[i], &c);
			*truncated += c;
		}
		n = njobs_sel;
	} else {
		double t = 0, lam = sp->rate_per_ms / ticks_per_ms, end = sp->duration_ms * ticks_per_ms;
		for (uint64_t j = 0;; j++) {
			t += -log(unit(smix(sp->seed, j, 1))) / lam;
			if (t >= end)
				break;
			if (n == cap) {
				cap *= 2;
				uint64_t *na = (uint64_t *)realloc(arr, cap * 8);
				if (!na)
					goto oom;
				arr = na;
				uint32_t *nl = (uint32_t *)realloc(len, cap * 4);
				if (!nl)
					goto oom;
				len = nl;
			}
			int c;
			arr[n] = (uint64_t)t;
			len[n] = job_length(sp, j, &c);
			*truncated += c;
			n++;
		}
	}
	*arr_out = arr;
	*len_out = len;
	*n_out = n;
	return 0;
oom:
	free(arr);
	free(len);
	set_err("out of host memory");
	return -1;
}

int64_t kma_serve_trace(const kma_serve_spec *spec, uint64_t *arrival, uint32_t *lengths, uint64_t max)
{
	if (!spec || spec->rate_per_ms <= 0 || spec->duration_ms <= 0 || spec->max_len == 0) {
		set_err("serving trace: bad spec");
		return -1;
	}
	uint64_t *arr = NULL, n = 0, tr = 0;
	uint32_t *len = NULL;
	if (gen_trace(spec, NULL, 0, 1e6 / kma_i_wall_clock_ns(), &arr, &len, &n, &tr) != 0)
		return -1;
	for (uint64_t j = 0; j < n && j < max; j++) {
		if (arrival)
			arrival[j] = arr[j];
		if (lengths)
			lengths[j] = len[j];
	}
	free(arr);
	free(len);
	return (int64_t)n;
}

int kma_serve_lengths(const kma_serve_spec *spec, const uint64_t *jobs, uint64_t n, uint32_t *lengths)
{
	if (!spec || (n && (!jobs || !lengths)) || spec->max_len == 0) {
		set_err("bad arguments");
		return -1;
	}
	for (uint64_t i = 0; i < n; i++) {
		int c;
		lengths[i] = job_length(spec, jobs[i], &c);
	}
	return 0;
}

static int cmp_u64(const void *a, const void *b)
{
	uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
	return x < y ? -1 : x > y;
}

static double hist_pct(const uint64_t *hist, uint32_t n, double q, double tick_ns)
{
	uint64_t tot = 0;
	for (uint32_t i = 0; i < n; i++)
		tot += hist[i];
	if (!tot)
		return 0;
	uint64_t want = (uint64_t)(q * (double)(tot - 1)), acc = 0;
	for (uint32_t i = 0; i < n; i++) {
		acc += hist[i];
		if (acc > want)
			return i * tick_ns;
	}
	return (n - 1) * tick_ns;
}

int kma_serve_run(kma_heap *h, const kma_serve_spec *spec, const uint64_t *jobs, uint64_t njobs_sel,
		  uint64_t *results, uint64_t nresults, kma_serve_stats *out)
{
	if (!h || !spec || !out || h->cfg.layout != KMA_LAYOUT_NATIVE || (jobs && njobs_sel == 0)) {
		set_err("serving loop: bad arguments (native heap required)");
		return -1;
	}
	kma_serve_spec sp = *spec;
	if (sp.max_admit == 0)
		sp.max_admit = 1024;
	if (sp.max_live == 0 || (sp.max_live & 255u) || sp.max_len == 0 || sp.block_bytes == 0 ||
	    sp.watermark_pct == 0 || sp.watermark_pct > (sp.admit == KMA_SERVE_ADMIT_PROJECTED ? 400u : 100u) ||
	    sp.duration_ms <= 0 || (!jobs && sp.rate_per_ms <= 0) || sp.admit > KMA_SERVE_ADMIT_WATERMARK) {
		set_err("serving loop: bad spec");
		return -1;
	}
	if (sp.admit == KMA_SERVE_ADMIT_PROJECTED && sp.max_len > KMA_SERVE_MAX_PROJECTED_LEN) {
		set_err("serving loop: projected admission looks at most %u steps ahead (max_len %u)",
			KMA_SERVE_MAX_PROJECTED_LEN, sp.max_len);
		return -1;
	}
	/* blocks of block_bytes the heap holds: superblocks x the slots of the
	 * smallest class that fits a block */
	uint64_t heap_blocks = 0;
	for (uint32_t c = 0; c < h->nh.nclasses; c++)
		if (h->nh.class_size[c] >= sp.block_bytes) {
			heap_blocks = h->nh.nsb * (uint64_t)h->nh.class_cap[c];
			break;
		}
	if (heap_blocks == 0) {
		set_err("serving loop: no size class holds %u-byte blocks", sp.block_bytes);
		return -1;
	}
	uint64_t cap_blocks = heap_blocks * sp.watermark_pct / 100u;
	if (cap_blocks > (1ull << 31))
		cap_blocks = 1ull << 31; /* timeline entries are u32 */
	memset(out, 0, sizeof(*out));
	double tick_ns = kma_i_wall_clock_ns(), ticks_per_ms = 1e6 / tick_ns;

	/* ---- job trace ---- */
	uint64_t *arr = NULL, n = 0;
	uint32_t *len = NULL;
	if (gen_trace(&sp, jobs, njobs_sel, ticks_per_ms, &arr, &len, &n, &out->truncated) != 0)
		return -1;
	out->jobs = n;

	/* ---- device buffers ---- */
	uint32_t maxg = 0;
	if (kmad_serve_max_groups(&maxg) != 0 || maxg == 0) {
		free(arr);
		free(len);
		set_err("serving loop: occupancy query failed");
		return -1;
	}
	/* default: one work-group per 256 job slots (the barriers' cost grows
	 * with the grid), at least 64, within the co-resident limit */
	uint32_t want = sp.max_live / 256u < 64u ? 64u : sp.max_live / 256u;
	uint32_t nwg = sp.nwg ? sp.nwg : (want < maxg ? want : maxg);
	if (nwg > maxg) {
		free(arr);
		free(len);
		set_err("serving loop: %u work-groups cannot all be resident (max %u)", nwg, maxg);
		return -1;
	}
	out->nwg = nwg;
	kma_serve_args a;
	memset(&a, 0, sizeof(a));
	a.heap = h->dev;
	a.njobs = n;
	a.max_live = sp.max_live;
	a.max_len = sp.max_len;
	a.block_bytes = sp.block_bytes;
	a.watermark_pct = sp.watermark_pct;
	a.max_admit = sp.max_admit;
	a.admit = sp.admit;
	a.cap_blocks = cap_blocks;
	a.keep_live = sp.keep_live != 0;
	a.duration_ticks = (uint64_t)(sp.duration_ms * ticks_per_ms);
	void *bufs[16] = {0};
	uint64_t nn = n ? n : 1, rq = (uint64_t)sp.max_live * 4;
	/* [5]: the free-slot stack, filled on the device */
	uint64_t sizes[16] = {nn * 8, nn * 4, (uint64_t)sp.max_live * 8, (uint64_t)sp.max_live * 4,
			      (uint64_t)sp.max_live * sp.max_len * 8, (uint64_t)sp.max_live * 4, KMA_SV_CTL_WORDS * 8,
			      (uint64_t)sp.max_live * 32, nn * 8, nn * 8, nn * 8, KMA_SV_NHIST * 8,
			      (uint64_t)sp.max_live * 4, (uint64_t)sp.max_live * 16, rq * 8, nn * 8};
	int rc = -1;
	hipStream_t st = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	uint64_t *hist = NULL, *adm = NULL, *waits = NULL;
	for (int i = 0; i < 16; i++)
		if ((i != 15 || jobs) && hipMalloc(&bufs[i], sizes[i]) != hipSuccess) {
			set_err("serving loop: device allocation of %llu bytes failed", (unsigned long long)sizes[i]);
			goto done;
		}
	a.arrival = (const uint64_t *)bufs[0];
	a.length = (const uint32_t *)bufs[1];
	a.slot_job = (int64_t *)bufs[2];
	a.slot_steps = (uint32_t *)bufs[3];
	a.table = (uint64_t *)bufs[4];
	a.free_slots = (uint32_t *)bufs[5];
	a.ctl = (uint64_t *)bufs[6];
	a.done = (uint64_t *)bufs[7];
	a.result = (uint64_t *)bufs[8];
	a.job_admit = (uint64_t *)bufs[9];
	a.job_done = (uint64_t *)bufs[10];
	a.lat_hist = (uint64_t *)bufs[11];
	a.slot_seq = (uint32_t *)bufs[12];
	a.lstack = (uint64_t *)bufs[13];
	a.lstack_cap = (uint64_t)sp.max_live * 2;
	a.requeue = (uint64_t *)bufs[14];
	a.rq_cap = rq;
	a.job_id = jobs ? (const uint64_t *)bufs[15] : NULL;
	uint64_t ctl0[KMA_SV_CTL_WORDS];
	memset(ctl0, 0, sizeof(ctl0));
	if ((jobs && hipMemcpy(bufs[15], jobs, n * 8, hipMemcpyHostToDevice) != hipSuccess) ||
	    hipMemset(bufs[12], 0, sizes[12]) != hipSuccess ||
	    (n && hipMemcpy(bufs[0], arr, n * 8, hipMemcpyHostToDevice) != hipSuccess) ||
	    (n && hipMemcpy(bufs[1], len, n * 4, hipMemcpyHostToDevice) != hipSuccess) ||
	    hipMemset(bufs[2], 0xff, sizes[2]) != hipSuccess || hipMemset(bufs[3], 0, sizes[3]) != hipSuccess ||
	    hipMemcpy(bufs[6], ctl0, sizeof(ctl0), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemset(bufs[8], 0, sizes[8]) != hipSuccess || hipMemset(bufs[9], 0, sizes[9]) != hipSuccess ||
	    hipMemset(bufs[10], 0, sizes[10]) != hipSuccess || hipMemset(bufs[11], 0, sizes[11]) != hipSuccess ||
	    hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
	    hipEventCreate(&e1) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
		set_err("serving loop: setup failed");
		goto done;
	}

	/* ---- run ---- */
	hipEventRecord(e0, st);
	if (kmad_serve_launch(&a, nwg, st) != 0) {
		set_err("serving loop: cooperative launch failed: %s", hipGetErrorString(hipGetLastError()));
		goto done;
	}
	hipEventRecord(e1, st);
	if (hipEventSynchronize(e1) != hipSuccess) {
		set_err("serving loop kernel failed: %s", hipGetErrorString(hipGetLastError()));
		goto done;
	}
	float ms = 0;
	hipEventElapsedTime(&ms, e0, e1);
	out->kernel_ms = ms;

	/* ---- statistics ---- */
	uint64_t ctl[KMA_SV_CTL_WORDS];
	hist = (uint64_t *)malloc(KMA_SV_NHIST * 8);
	if (!hist || hipMemcpy(ctl, bufs[6], sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess ||
	    hipMemcpy(hist, bufs[11], KMA_SV_NHIST * 8, hipMemcpyDeviceToHost) != hipSuccess) {
		set_err("serving loop: download failed");
		goto done;
	}
	const uint64_t *S = ctl + KMA_SV_STATS;
	out->aborted = ctl[SV_C_ABORT] != 0;
	out->steps = S[0];
	out->mallocs = S[1];
	out->stalls = S[2];
	out->frees = S[3];
	out->bad = S[4];
	out->admitted = S[5];
	out->done = S[6];
	out->peak_used_frac = h->nh.nsb ? (double)S[7] / (double)h->nh.nsb : 0;
	out->run_ms = (double)S[8] / ticks_per_ms;
	out->drained = S[9];
	out->drain_frees = S[10];
	out->preempted = S[11];
	out->rebuilds = S[12];
	out->proj_checked = S[21];
	out->proj_mismatch = S[22];
	out->admit = sp.admit;
	out->cap_blocks = cap_blocks;
	out->peak_proj_frac = (double)S[13] / (double)heap_blocks;
	out->mean_live = out->steps ? (double)S[14] / (double)out->steps : 0;
	{
		/* phase profile: S[15..18] work-group 0's wall clock per phase over
		 * steps + 1 phase-A passes (the last one decides to stop) and steps
		 * B and C phases; S[19], S[20] all groups' working time in B, C */
		double na = (double)(out->steps + 1u), nb = out->steps ? (double)out->steps : 1.0;
		out->phase_us[0] = S[15] * tick_ns * 1e-3 / na;
		out->phase_us[1] = S[16] * tick_ns * 1e-3 / na;
		out->phase_us[2] = S[17] * tick_ns * 1e-3 / nb;
		out->phase_us[3] = S[18] * tick_ns * 1e-3 / nb;
		out->busy_frac[0] = S[17] ? (double)S[19] / ((double)S[17] * nwg) : 0;
		out->busy_frac[1] = S[18] ? (double)S[20] / ((double)S[18] * nwg) : 0;
	}
	out->ops_per_s = out->run_ms > 0 ? (double)(out->mallocs + out->frees) / (out->run_ms * 1e-3) : 0;
	out->goodput_per_s = out->run_ms > 0 ? (double)out->done / (out->run_ms * 1e-3) : 0;
	out->preempt_frac = out->admitted ? (double)out->preempted / (double)out->admitted : 0;
	out->lat_p50_ns = hist_pct(hist, KMA_SV_NHIST, 0.50, tick_ns);
	out->lat_p99_ns = hist_pct(hist, KMA_SV_NHIST, 0.99, tick_ns);
	out->lat_max_ns = hist_pct(hist, KMA_SV_NHIST, 1.0, tick_ns);
	uint64_t nadm = ctl[SV_C_NEXT];
	if (nadm) {
		adm = (uint64_t *)malloc(nadm * 8);
		waits = (uint64_t *)malloc(nadm * 8);
		if (!adm || !waits || hipMemcpy(adm, bufs[9], nadm * 8, hipMemcpyDeviceToHost) != hipSuccess) {
			set_err("serving loop: download failed");
			goto done;
		}
		for (uint64_t j = 0; j < nadm; j++)
			waits[j] = adm[j] ? adm[j] - 1u - arr[j] : 0u;
		qsort(waits, nadm, 8, cmp_u64);
		out->wait_p50_us = waits[nadm / 2] * tick_ns * 1e-3;
		out->wait_p99_us = waits[(nadm * 99) / 100] * tick_ns * 1e-3;
	}
	if (results && nresults) {
		uint64_t k = nresults < n ? nresults : n;
		if (hipMemcpy(results, bufs[8], k * 8, hipMemcpyDeviceToHost) != hipSuccess) {
			set_err("serving loop: download failed");
			goto done;
		}
		uint64_t *dn = (uint64_t *)malloc(k * 8);
		if (!dn || hipMemcpy(dn, bufs[10], k * 8, hipMemcpyDeviceToHost) != hipSuccess) {
			free(dn);
			set_err("serving loop: download failed");
			goto done;
		}
		for (uint64_t j = 0; j < k; j++)
			if (!dn[j])
				results[j] = 0; /* not completed: partial sums of drained jobs are dropped */
		free(dn);
	}
	if (out->aborted) {
		set_err("serving loop: a grid barrier timed out; the loop was aborted");
		goto done;
	}
	rc = 0;
done:
	free(arr);
	free(len);
	free(hist);
	free(adm);
	free(waits);
	if (e0)
		hipEventDestroy(e0);
	if (e1)
		hipEventDestroy(e1);
	if (st)
		hipStreamDestroy(st);
	for (int i = 0; i < 16; i++)
		if (bufs[i])
			hipFree(bufs[i]);
	return rc;
}
