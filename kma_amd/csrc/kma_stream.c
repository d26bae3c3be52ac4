/*
 * kma_stream.c - batched request streams: host-side trace generation, the
 * run loop (one launch per batch, or one graph-captured persistent launch)
 * and the stream autotuner.  C over the HIP runtime C API.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kma.h"
#include "kma_internal.h"
#include "kma_launch.h"
#include "kma_work.h"

struct kma_stream {
	kma_stream_spec spec;
	kma_stream_info info;
	uint32_t *req;  /* host trace, info.nreq */
	uint32_t *boff; /* batch offsets, info.nbatch + 1 */
	uint32_t *d_req;
	uint64_t *d_res;
	uint64_t *d_ctl;
	uint64_t *d_ts; /* allocated by the first KMA_VAR_LOG run */
};

/* splitmix64 finaliser over a linear combination of (seed, a, b) */
static uint64_t smix(uint64_t seed, uint64_t a, uint64_t b)
{
	uint64_t z = seed * 0x9E3779B97F4A7C15ull + a * 0xBF58476D1CE4E5B9ull + b * 0x94D049BB133111EBull;
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static int log2_exact(uint32_t x)
{
	if (x == 0 || (x & (x - 1)))
		return -1;
	int e = 0;
	while ((1u << e) < x)
		e++;
	return e;
}

static uint64_t pow2_up(uint64_t x)
{
	uint64_t p = 1;
	while (p < x)
		p <<= 1;
	return p;
}

typedef struct u32vec {
	uint32_t *v;
	uint64_t n, cap;
} u32vec;

static int push(u32vec *a, uint32_t x)
{
	if (a->n == a->cap) {
		uint64_t nc = a->cap ? 2 * a->cap : 4096;
		uint32_t *nv = (uint32_t *)realloc(a->v, nc * 4);
		if (!nv)
			return -1;
		a->v = nv;
		a->cap = nc;
	}
	a->v[a->n++] = x;
	return 0;
}

/* Block footprint used for the demand figures in kma_stream_info. */
static void footprint(const kma_stream_spec *sp, uint32_t size, uint64_t *small, uint64_t *large)
{
	if (size >= sp->page_min)
		*large += ((uint64_t)size + sp->page_min - 1) / sp->page_min * sp->page_min;
	else
		*small += pow2_up(size);
}

/* The generator.  Batch b has Zipf(zipf_milli/1000) size over 1..max_batch,
 * drawn by inverse CDF from u = smix(seed, b, 1) / 2^64; request j is a free
 * when smix(seed, j, 2) % 100 < free_pct and a block of an earlier batch is
 * live: it frees live[smix(seed, j, 3) % nlive] (swap-remove); otherwise it
 * is a malloc of 2^e + (x >> 16) % 2^e bytes with x = smix(seed, j, 4) and e
 * uniform in [log2 min_size, log2 max_size) from x % (range). */
static int generate(kma_stream *s)
{
	const kma_stream_spec *sp = &s->spec;
	int elo = log2_exact(sp->min_size), ehi = log2_exact(sp->max_size);
	if (sp->nbatch == 0 || sp->max_batch == 0 || sp->free_pct > 100 || elo < 0 || ehi <= elo || ehi > 31 ||
	    sp->zipf_milli == 0 || sp->page_min == 0) {
		set_err("bad stream spec");
		return -1;
	}
	double *cdf = (double *)malloc((size_t)sp->max_batch * sizeof(double));
	u32vec req = {0, 0, 0}, live = {0, 0, 0}, pend = {0, 0, 0};
	uint32_t *boff = (uint32_t *)malloc(((size_t)sp->nbatch + 2) * 4);
	int rc = -1;
	if (!cdf || !boff) {
		set_err("out of host memory");
		goto out;
	}
	double zs = sp->zipf_milli / 1000.0, acc = 0;
	for (uint32_t i = 0; i < sp->max_batch; i++) {
		acc += pow((double)(i + 1), -zs);
		cdf[i] = acc;
	}
	for (uint32_t i = 0; i < sp->max_batch; i++)
		cdf[i] /= acc;
	cdf[sp->max_batch - 1] = 1.0;

	kma_stream_info *in = &s->info;
	memset(in, 0, sizeof(*in));
	uint64_t live_small = 0, live_large = 0;
	uint32_t nb = 0;
	for (uint32_t b = 0; b < sp->nbatch; b++) {
		double u = (double)(smix(sp->seed, b, 1) >> 11) * (1.0 / 9007199254740992.0);
		uint32_t lo = 0, hi = sp->max_batch - 1;
		while (lo < hi) {
			uint32_t mid = lo + (hi - lo) / 2;
			if (cdf[mid] > u)
				hi = mid;
			else
				lo = mid + 1;
		}
		uint64_t n = (uint64_t)lo + 1;
		if (n > in->max_batch)
			in->max_batch = n;
		boff[nb++] = (uint32_t)req.n;
		pend.n = 0;
		for (uint64_t i = 0; i < n; i++) {
			uint64_t j = req.n;
			if (j >= 0x7fffffffull) {
				set_err("stream too long");
				goto out;
			}
			if (live.n && smix(sp->seed, j, 2) % 100 < sp->free_pct) {
				uint64_t k = smix(sp->seed, j, 3) % live.n;
				uint32_t t = live.v[k];
				live.v[k] = live.v[--live.n];
				uint64_t a = 0, c = 0;
				footprint(sp, req.v[t], &a, &c);
				live_small -= a;
				live_large -= c;
				if (push(&req, 0x80000000u | t))
					goto oom;
				in->nfree++;
			} else {
				uint64_t x = smix(sp->seed, j, 4);
				uint32_t e = (uint32_t)elo + (uint32_t)(x % (uint64_t)(ehi - elo));
				uint32_t size = (1u << e) + (uint32_t)((x >> 16) % (1ull << e));
				if (push(&req, size) || push(&pend, (uint32_t)j))
					goto oom;
				footprint(sp, size, &in->small_bytes, &in->large_bytes);
				in->nmalloc++;
			}
		}
		for (uint64_t i = 0; i < pend.n; i++) {
			if (push(&live, pend.v[i]))
				goto oom;
			footprint(sp, req.v[pend.v[i]], &live_small, &live_large);
		}
		if (live.n > in->peak_live)
			in->peak_live = live.n;
		if (live_small > in->peak_small_bytes)
			in->peak_small_bytes = live_small;
		if (live_large > in->peak_large_bytes)
			in->peak_large_bytes = live_large;
	}
	if (sp->drain && live.n) {
		boff[nb++] = (uint32_t)req.n;
		for (uint64_t i = 0; i < live.n; i++) {
			if (push(&req, 0x80000000u | live.v[i]))
				goto oom;
			in->nfree++;
		}
	}
	boff[nb] = (uint32_t)req.n;
	in->nbatch = nb;
	in->nreq = req.n;
	s->req = req.v;
	req.v = NULL;
	s->boff = boff;
	boff = NULL;
	rc = 0;
	goto out;
oom:
	set_err("out of host memory");
out:
	free(cdf);
	free(req.v);
	free(live.v);
	free(pend.v);
	free(boff);
	return rc;
}

kma_stream *kma_stream_create(const kma_stream_spec *spec)
{
	if (!spec) {
		set_err("NULL spec");
		return NULL;
	}
	kma_stream *s = (kma_stream *)calloc(1, sizeof(*s));
	if (!s)
		return NULL;
	s->spec = *spec;
	if (generate(s) != 0) {
		free(s);
		return NULL;
	}
	return s;
}

/* Device copy, made by the first run (creation needs no device). */
static int upload(kma_stream *s)
{
	if (s->d_req)
		return 0;
	uint64_t n = s->info.nreq;
	if (hipMalloc((void **)&s->d_req, (n ? n : 1) * 4) != hipSuccess ||
	    hipMalloc((void **)&s->d_res, (n ? n : 1) * 8) != hipSuccess ||
	    hipMalloc((void **)&s->d_ctl, KMA_CTL_WORDS * 8) != hipSuccess ||
	    (n && hipMemcpy(s->d_req, s->req, n * 4, hipMemcpyHostToDevice) != hipSuccess) ||
	    hipMemset(s->d_res, 0, (n ? n : 1) * 8) != hipSuccess ||
	    hipMemset(s->d_ctl, 0, KMA_CTL_WORDS * 8) != hipSuccess) {
		set_err("stream upload failed");
		return -1;
	}
	return 0;
}

void kma_stream_destroy(kma_stream *s)
{
	if (!s)
		return;
	if (s->d_req)
		hipFree(s->d_req);
	if (s->d_res)
		hipFree(s->d_res);
	if (s->d_ctl)
		hipFree(s->d_ctl);
	if (s->d_ts)
		hipFree(s->d_ts);
	free(s->req);
	free(s->boff);
	free(s);
}

int kma_stream_get_info(const kma_stream *s, kma_stream_info *out)
{
	if (!s || !out)
		return -1;
	*out = s->info;
	return 0;
}

int kma_stream_trace(const kma_stream *s, uint32_t *req, uint32_t *batch_off)
{
	if (!s)
		return -1;
	if (req)
		memcpy(req, s->req, s->info.nreq * 4);
	if (batch_off)
		memcpy(batch_off, s->boff, (s->info.nbatch + 1) * 4);
	return 0;
}

int kma_stream_results(const kma_stream *s, uint64_t *res)
{
	if (!s || !res || !s->d_res) {
		set_err("no run yet");
		return -1;
	}
	HIPCHK(hipMemcpy(res, s->d_res, s->info.nreq * 8, hipMemcpyDeviceToHost), -1);
	for (uint64_t j = 0; j < s->info.nreq; j++)
		res[j] = (s->req[j] >> 31) ? 0 : (res[j] & ((1ull << 40) - 1ull));
	return 0;
}

int kma_stream_times(const kma_stream *s, uint64_t *ticks)
{
	if (!s || !ticks || !s->d_ts) {
		set_err("no logged run");
		return -1;
	}
	HIPCHK(hipMemcpy(ticks, s->d_ts, s->info.nreq * 8, hipMemcpyDeviceToHost), -1);
	return 0;
}

/* One replay: the begin kernel, then the form's launches. */
static int enqueue_replay(kma_heap *h, kma_stream *s, uint32_t form, uint32_t nwg, uint32_t per_thread,
			  uint32_t variant, hipStream_t st, uint64_t *launches)
{
	int rc = kmad_stream_begin(s->d_ctl, st);
	if (rc)
		return rc;
	if (form == KMA_FORM_BATCHED) {
		*launches = 1;
		return kmad_stream_persistent(h->dev, s->d_req, s->d_res, s->info.nreq, nwg, per_thread, s->d_ctl,
					      s->d_ts, variant, st);
	}
	*launches = 0;
	for (uint64_t b = 0; b < s->info.nbatch; b++) {
		uint64_t lo = s->boff[b], n = (uint64_t)s->boff[b + 1] - lo;
		if (!n)
			continue;
		rc = kmad_stream_plain(h->dev, s->d_req, s->d_res, lo, n, s->d_ctl, s->d_ts, variant, st);
		if (rc)
			return rc;
		++*launches;
	}
	return 0;
}

int kma_stream_run(kma_heap *h, kma_stream *s, uint32_t form, uint32_t nwg, uint32_t per_thread, uint32_t warmup,
		   uint32_t reps, uint32_t variant, kma_stream_stats *out)
{
	if (!h || !s || !out || reps == 0) {
		set_err("bad arguments");
		return -1;
	}
	memset(out, 0, sizeof(*out));
	if (h->cfg.layout != KMA_LAYOUT_NATIVE) {
		set_err("streams run on a native-layout heap");
		return -1;
	}
	if (form != KMA_FORM_PLAIN && form != KMA_FORM_BATCHED) {
		set_err("stream form must be PLAIN or BATCHED");
		return -1;
	}
	if (form == KMA_FORM_BATCHED && (nwg == 0 || nwg > 65536 || per_thread == 0 || per_thread > 4096)) {
		set_err("bad persistent launch shape");
		return -1;
	}
	if (upload(s) != 0)
		return -1;
	if ((variant & KMA_VAR_LOG) && !s->d_ts &&
	    hipMalloc((void **)&s->d_ts, (s->info.nreq ? s->info.nreq : 1) * 8) != hipSuccess) {
		set_err("log buffer allocation failed");
		return -1;
	}
	hipStream_t st = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	hipGraph_t graph = NULL;
	hipGraphExec_t gexec = NULL;
	uint64_t st0[NAT_NSTATS], st1[NAT_NSTATS], ctl[KMA_CTL_WORDS];
	uint64_t launches = 0;
	int rc = -1;
	if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
	    hipEventCreate(&e1) != hipSuccess) {
		set_err("stream setup failed");
		goto done;
	}
	if (form == KMA_FORM_BATCHED) {
		/* the whole replay (begin + persistent kernel) is one graph */
		if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
			set_err("graph capture failed");
			goto done;
		}
		int lr = enqueue_replay(h, s, form, nwg, per_thread, variant, st, &launches);
		if (hipStreamEndCapture(st, &graph) != hipSuccess || lr != 0 ||
		    hipGraphInstantiate(&gexec, graph, NULL, NULL, 0) != hipSuccess) {
			set_err("graph build failed: %d", lr);
			goto done;
		}
	}
	for (uint32_t r = 0; r < warmup + reps; r++) {
		if (r == warmup) {
			if (hipStreamSynchronize(st) != hipSuccess || kma_i_read_native_stats(h, st0) != 0) {
				set_err("warm-up failed: %s", hipGetErrorString(hipGetLastError()));
				goto done;
			}
			hipEventRecord(e0, st);
		}
		if (gexec) {
			if (hipGraphLaunch(gexec, st) != hipSuccess) {
				set_err("graph launch failed");
				goto done;
			}
		} else {
			int lr = enqueue_replay(h, s, form, nwg, per_thread, variant, st, &launches);
			if (lr) {
				set_err("launch failed: %d", lr);
				goto done;
			}
		}
	}
	hipEventRecord(e1, st);
	if (hipEventSynchronize(e1) != hipSuccess) {
		set_err("stream kernels failed: %s", hipGetErrorString(hipGetLastError()));
		goto done;
	}
	float ms = 0;
	hipEventElapsedTime(&ms, e0, e1);
	if (hipMemcpy(ctl, s->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost) != hipSuccess ||
	    kma_i_read_native_stats(h, st1) != 0) {
		set_err("stats download failed");
		goto done;
	}
	out->gpu_ms = ms / reps;
	out->launches = launches;
	out->kernel_ms_avg = launches ? out->gpu_ms / (double)launches : 0;
	{
		const uint64_t *c = ctl + KMA_CTL_STATS;
		out->ops = c[0];
		out->mallocs = c[1];
		out->nulls = c[2];
		out->frees = c[3];
		out->free_nulls = c[4];
		out->bad = c[5];
		out->wait_timeouts = c[6];
		out->large = c[7];
		out->large_nulls = c[8];
		if (c[12])
			for (int k = 0; k < 3; k++)
				out->phase_ns[k] = kma_i_wall_clock_ns() * (double)c[9 + k] / (double)c[12];
	}
	out->double_frees = (st1[NAT_STAT_DOUBLE_FREE] - st0[NAT_STAT_DOUBLE_FREE]) / reps;
	out->bad_frees = (st1[NAT_STAT_BAD_FREE] - st0[NAT_STAT_BAD_FREE]) / reps;
	out->pool_empty = (st1[2] - st0[2]) / reps;
	for (int k = 0; k < 4; k++)
		out->waits[k] = (st1[3 + k] - st0[3 + k]) / reps;
	out->remote_frees = (st1[7] - st0[7]) / reps;
	out->listed = (st1[9] - st0[9]) / reps;
	out->reattached = (st1[10] - st0[10]) / reps;
	out->tick_ns = kma_i_wall_clock_ns();
	rc = 0;
done:
	if (gexec)
		hipGraphExecDestroy(gexec);
	if (graph)
		hipGraphDestroy(graph);
	if (e0)
		hipEventDestroy(e0);
	if (e1)
		hipEventDestroy(e1);
	if (st)
		hipStreamDestroy(st);
	return rc;
}

int kma_stream_heap_config(const kma_stream *s, uint64_t heap_bytes, uint32_t sb_log2, uint32_t table,
			   uint32_t page_log2, kma_config *out)
{
	if (!s || !out || sb_log2 < 12 || sb_log2 > 20 || page_log2 < 12 || page_log2 > 30 ||
	    table > KMA_TABLE_HALF) {
		set_err("bad heap request");
		return -1;
	}
	memset(out, 0, sizeof(*out));
	out->layout = KMA_LAYOUT_NATIVE;
	out->sb_log2 = sb_log2;
	out->page_log2 = page_log2;
	uint64_t top = 1ull << sb_log2;
	if (s->spec.page_min < top)
		top = s->spec.page_min;
	uint32_t n = 0;
	for (uint64_t c = s->spec.min_size < 8 ? 8 : s->spec.min_size; c <= top && n < KMA_MAX_CLASSES; c <<= 1) {
		out->class_size[n++] = (uint32_t)c;
		if (table == KMA_TABLE_HALF && c + c / 2 <= top && c >= 16 && n < KMA_MAX_CLASSES)
			out->class_size[n++] = (uint32_t)(c + c / 2);
	}
	out->nclasses = n;
	double dem = (double)s->info.small_bytes + (double)s->info.large_bytes;
	double fs = dem > 0 ? (double)s->info.small_bytes / dem : 1.0;
	uint64_t small = (uint64_t)(fs * (double)heap_bytes);
	out->sblocks = small >> sb_log2;
	out->npages = (heap_bytes - (out->sblocks << sb_log2)) >> page_log2;
	if (s->info.large_bytes == 0)
		out->npages = 0;
	if (out->sblocks == 0 && out->npages == 0) {
		set_err("heap too small");
		return -1;
	}
	return 0;
}

int kma_stream_autotune(kma_stream *s, uint64_t heap_bytes, uint32_t page_log2, uint64_t null_limit, uint32_t reps,
			kma_tune_result *out)
{
	static const uint32_t sbs[] = {14, 15, 16};
	static const uint32_t aggs[] = {KMA_VAR_AGG_WAVE, KMA_VAR_AGG_WG8, KMA_VAR_AGG_WG16, KMA_VAR_AGG_WG32};
	static const uint32_t nwgs[] = {512, 1024};
	static const uint32_t pts[] = {1, 2};
	if (!s || !out || reps == 0) {
		set_err("bad arguments");
		return -1;
	}
	memset(out, 0, sizeof(*out));
	int best = -1, fewest = -1;
	for (uint32_t a = 0; a < sizeof(sbs) / sizeof(sbs[0]); a++) {
		for (uint32_t table = KMA_TABLE_POW2; table <= KMA_TABLE_HALF; table++) {
			kma_config cfg;
			if (kma_stream_heap_config(s, heap_bytes, sbs[a], table, page_log2, &cfg) != 0)
				return -1;
			kma_heap *h = kma_create_ex(&cfg);
			if (!h)
				return -1;
			for (uint32_t g = 0; g < sizeof(aggs) / sizeof(aggs[0]); g++) {
			for (uint32_t b = 0; b < sizeof(nwgs) / sizeof(nwgs[0]); b++) {
				for (uint32_t c = 0; c < sizeof(pts) / sizeof(pts[0]); c++) {
					kma_stream_stats st;
					if (out->ntrials >= KMA_TUNE_MAX ||
					    kma_stream_run(h, s, KMA_FORM_BATCHED, nwgs[b], pts[c], 1, reps, aggs[g], &st) != 0) {
						kma_destroy(h);
						return -1;
					}
					kma_tune_trial *t = &out->trials[out->ntrials];
					t->sb_log2 = sbs[a];
					t->table = table;
					t->nwg = nwgs[b];
					t->per_thread = pts[c];
					t->agg = aggs[g];
					t->ms = st.gpu_ms;
					t->ops_per_s = st.gpu_ms > 0 ? (double)st.ops / (st.gpu_ms * 1e-3) : 0;
					t->nulls = st.nulls;
					t->violations = st.bad + st.wait_timeouts + st.double_frees + st.bad_frees +
							st.waits[0] + st.waits[1] + st.waits[2] + st.waits[3];
					int i = (int)out->ntrials++;
					if (t->violations)
						continue;
					if (fewest < 0 || t->nulls < out->trials[fewest].nulls)
						fewest = i;
					if ((null_limit == 0 || t->nulls <= null_limit) &&
					    (best < 0 || t->ops_per_s > out->trials[best].ops_per_s))
						best = i;
				}
			}
			}
			kma_destroy(h);
		}
	}
	if (best < 0)
		best = fewest;
	if (best < 0) {
		set_err("every trial had violations");
		return -1;
	}
	/* last trial: the chosen configuration without reuse of partly free
	 * superblocks (on a trace that keeps the heap oversubscribed the reuse
	 * can cost more refills than the capacity it recovers) */
	{
		const kma_tune_trial b = out->trials[best];
		kma_config cfg;
		if (kma_stream_heap_config(s, heap_bytes, b.sb_log2, b.table, page_log2, &cfg) != 0)
			return -1;
		cfg.no_partial = 1;
		kma_heap *h = kma_create_ex(&cfg);
		if (!h)
			return -1;
		kma_stream_stats st;
		if (out->ntrials >= KMA_TUNE_MAX ||
		    kma_stream_run(h, s, KMA_FORM_BATCHED, b.nwg, b.per_thread, 1, reps, b.agg, &st) != 0) {
			kma_destroy(h);
			return -1;
		}
		kma_destroy(h);
		kma_tune_trial *t = &out->trials[out->ntrials];
		*t = b;
		t->no_partial = 1;
		t->ms = st.gpu_ms;
		t->ops_per_s = st.gpu_ms > 0 ? (double)st.ops / (st.gpu_ms * 1e-3) : 0;
		t->nulls = st.nulls;
		t->violations = st.bad + st.wait_timeouts + st.double_frees + st.bad_frees + st.waits[0] +
				st.waits[1] + st.waits[2] + st.waits[3];
		const int i = (int)out->ntrials++;
		if (!t->violations && (null_limit == 0 || t->nulls <= null_limit) &&
		    t->ops_per_s > out->trials[best].ops_per_s)
			best = i;
	}
	const kma_tune_trial *t = &out->trials[best];
	if (kma_stream_heap_config(s, heap_bytes, t->sb_log2, t->table, page_log2, &out->cfg) != 0)
		return -1;
	out->cfg.no_partial = t->no_partial;
	out->nwg = t->nwg;
	out->per_thread = t->per_thread;
	out->table = t->table;
	out->agg = t->agg;
	out->ops_per_s = t->ops_per_s;
	out->nulls = t->nulls;
	return 0;
}
