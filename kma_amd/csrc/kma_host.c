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
#include "kma_work.h"

/* Must match kma_dheap in kma_device.h (checked by kma_native_header_size). */
typedef struct nat_hdr {
	uint64_t magic;
	uint64_t bytes;
	uint64_t nsb;
	uint64_t npages;
	uint32_t sb_log2;
	uint32_t page_log2;
	uint32_t nclasses;
	uint32_t nctx;
	uint32_t npool;
	uint32_t bm_words;
	uint32_t max_small;
	uint32_t pad0;
	uint64_t off_state;
	uint64_t off_next;
	uint64_t off_bitmap;
	uint64_t off_heads;
	uint64_t off_pool;
	uint64_t off_pbitmap;
	uint64_t off_plen;
	uint64_t off_data;
	uint64_t off_pages;
	uint64_t off_stats;
	uint32_t class_size[32];
	uint32_t class_cap[32];
	uint64_t class_magic[32];
	uint8_t class_for_log2[64];
} nat_hdr;

#define NAT_MAGIC 0x4e414d4b31414d4bull
#define NAT_HEADS_PER_CTX 32u
#define NAT_POOL_STRIDE_BYTES 128u
#define NAT_NSTATS 8u
#define NAT_STAT_DOUBLE_FREE 0
#define NAT_STAT_BAD_FREE 1

struct kma_heap {
	kma_config cfg;
	char *dev;        /* device buffer == header address */
	uint64_t bytes;   /* bytes allocated */
	nat_hdr nh;       /* native header (host copy) */
};

static __thread char g_err[256];

static void set_err(const char *fmt, ...)
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

#define HIPCHK(call, ret)                                                            \
	do {                                                                         \
		hipError_t e_ = (call);                                              \
		if (e_ != hipSuccess) {                                              \
			set_err("%s failed: %s", #call, hipGetErrorString(e_));      \
			return ret;                                                  \
		}                                                                    \
	} while (0)

#define LCHK(call, ret)                                                              \
	do {                                                                         \
		int r_ = (call);                                                     \
		if (r_ != 0) {                                                       \
			set_err("%s failed: %d (%s)", #call, r_,                     \
				hipGetErrorString((hipError_t)(-r_)));               \
			return ret;                                                  \
		}                                                                    \
	} while (0)

int kma_device_ok(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
		set_err("no HIP device");
		return -1;
	}
	return 0;
}

static uint64_t align_up(uint64_t x, uint64_t a)
{
	return (x + a - 1) & ~(a - 1);
}

static int is_pow2(uint64_t x)
{
	return x && !(x & (x - 1));
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
	if (hipMalloc((void **)&h->dev, h->bytes) != hipSuccess) {
		set_err("KMA: Could not allocate heap on-device (%llu bytes)", (unsigned long long)h->bytes);
		free(h);
		return NULL;
	}
	if (kma_reset(h) != 0) {
		hipFree(h->dev);
		free(h);
		return NULL;
	}
	return h;
}

/* ---------------- native layout ---------------- */
static int native_layout(kma_heap *h, const kma_config *in)
{
	kma_config *cfg = &h->cfg;
	*cfg = *in;
	if (cfg->sb_log2 == 0)
		cfg->sb_log2 = 14;
	if (cfg->sb_log2 < 10 || cfg->sb_log2 > 20) {
		set_err("native layout: sb_log2 %u outside 10..20", cfg->sb_log2);
		return -1;
	}
	uint32_t sbsz = 1u << cfg->sb_log2;
	if (cfg->nclasses == 0) {
		uint32_t n = 0;
		for (uint32_t s = 8; s <= sbsz / 2 && n < KMA_MAX_CLASSES; s <<= 1)
			cfg->class_size[n++] = s;
		cfg->nclasses = n;
	}
	if (cfg->nclasses > KMA_MAX_CLASSES) {
		set_err("native layout: %u classes > %u", cfg->nclasses, KMA_MAX_CLASSES);
		return -1;
	}
	for (uint32_t c = 0; c < cfg->nclasses; c++) {
		uint32_t s = cfg->class_size[c];
		if (s < 8 || (s & 7) || s > sbsz || (c && s <= cfg->class_size[c - 1])) {
			set_err("native layout: bad class table entry %u (%u B)", c, s);
			return -1;
		}
	}
	if (cfg->sblocks == 0 && cfg->npages == 0) {
		set_err("native layout: empty heap");
		return -1;
	}
	if (cfg->sblocks > 0xfffffff0ull) {
		set_err("native layout: too many superblocks");
		return -1;
	}
	if (cfg->nctx == 0) {
		uint64_t want = cfg->sblocks / (4ull * cfg->nclasses);
		uint32_t n = 1;
		while ((uint64_t)n * 2 <= want && n < 1024)
			n *= 2;
		cfg->nctx = n;
	}
	if (!is_pow2(cfg->nctx) || cfg->nctx > 65536) {
		set_err("native layout: nctx %u must be a power of two <= 65536", cfg->nctx);
		return -1;
	}
	if (cfg->npool == 0)
		cfg->npool = cfg->nctx < 256 ? cfg->nctx : 256;
	if (!is_pow2(cfg->npool) || cfg->npool > 65536) {
		set_err("native layout: npool %u must be a power of two <= 65536", cfg->npool);
		return -1;
	}
	if (cfg->page_log2 == 0)
		cfg->page_log2 = cfg->sb_log2;
	if (cfg->page_log2 < 12 || cfg->page_log2 > 30) {
		set_err("native layout: page_log2 %u outside 12..30", cfg->page_log2);
		return -1;
	}

	nat_hdr *nh = &h->nh;
	memset(nh, 0, sizeof(*nh));
	nh->magic = NAT_MAGIC;
	nh->nsb = cfg->sblocks;
	nh->npages = cfg->npages;
	nh->sb_log2 = cfg->sb_log2;
	nh->page_log2 = cfg->page_log2;
	nh->nclasses = cfg->nclasses;
	nh->nctx = cfg->nctx;
	nh->npool = cfg->npool;
	uint32_t maxcap = sbsz / cfg->class_size[0];
	nh->bm_words = (maxcap + 31) / 32;
	nh->max_small = cfg->class_size[cfg->nclasses - 1];
	for (uint32_t c = 0; c < cfg->nclasses; c++) {
		uint32_t s = cfg->class_size[c];
		nh->class_size[c] = s;
		nh->class_cap[c] = sbsz / s;
		nh->class_magic[c] = ((1ull << 40) + s - 1) / s;
	}
	for (uint32_t c = cfg->nclasses; c < 32; c++)
		nh->class_size[c] = 0xffffffffu;
	for (uint32_t e = 0; e < 64; e++) {
		/* first class that can hold 2^(e-1)+1 bytes */
		uint64_t lo = e == 0 ? 1 : (1ull << (e - 1)) + 1;
		uint32_t c = 0;
		while (c < cfg->nclasses && cfg->class_size[c] < lo)
			c++;
		if (c >= cfg->nclasses)
			c = cfg->nclasses - 1;
		nh->class_for_log2[e] = (uint8_t)c;
	}
	uint64_t off = align_up(sizeof(nat_hdr), 256);
	nh->off_stats = off;
	off += 256;
	nh->off_state = off;
	off = align_up(off + nh->nsb * 8, 256);
	nh->off_next = off;
	off = align_up(off + nh->nsb * 4, 256);
	nh->off_bitmap = off;
	off = align_up(off + nh->nsb * nh->bm_words * 4ull, 256);
	nh->off_heads = off;
	off = align_up(off + (uint64_t)nh->nctx * NAT_HEADS_PER_CTX * 8, 256);
	nh->off_pool = off;
	off = align_up(off + (uint64_t)nh->npool * NAT_POOL_STRIDE_BYTES, 256);
	nh->off_pbitmap = off;
	off = align_up(off + ((nh->npages + 63) / 64) * 8, 256);
	nh->off_plen = off;
	off = align_up(off + nh->npages, 256);
	/* metadata bytes; data placed after allocation so it is absolutely
	 * aligned to the superblock and page sizes */
	nh->off_data = off;
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
	uint64_t sbsz = 1ull << nh->sb_log2, psz = 1ull << nh->page_log2;
	uint64_t meta = nh->off_data;
	uint64_t total = meta + sbsz + (nh->nsb << nh->sb_log2) + psz + (nh->npages << nh->page_log2);
	if (hipMalloc((void **)&h->dev, total) != hipSuccess) {
		set_err("KMA: Could not allocate heap on-device (%llu bytes)", (unsigned long long)total);
		free(h);
		return NULL;
	}
	h->bytes = total;
	uint64_t base = (uint64_t)(uintptr_t)h->dev;
	nh->off_data = align_up(base + meta, sbsz) - base;
	nh->off_pages = align_up(base + nh->off_data + (nh->nsb << nh->sb_log2), psz) - base;
	nh->bytes = total;
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
		 * image this library defines starts from zero. */
		HIPCHK(hipMemset(h->dev, 0, h->bytes), -1);
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

int kma_native_header_size(void)
{
	return (int)sizeof(nat_hdr);
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
	LCHK(kmad_trim(h->dev, h->nh.nctx, h->nh.nclasses), -1);
	return 0;
}

static int read_native_stats(kma_heap *h, uint64_t *st)
{
	if (h->cfg.layout != KMA_LAYOUT_NATIVE) {
		memset(st, 0, NAT_NSTATS * 8);
		return 0;
	}
	HIPCHK(hipMemcpy(st, h->dev + h->nh.off_stats, NAT_NSTATS * 8, hipMemcpyDeviceToHost), -1);
	return 0;
}

int kma_run_workload(kma_heap *h, const kma_workload *wl, kma_run_stats *out)
{
	if (!h || !wl || !out) {
		set_err("bad arguments");
		return -1;
	}
	memset(out, 0, sizeof(*out));
	if (wl->mode > KMA_WL_HIGHVAR || wl->form > KMA_FORM_FUSED || wl->nwi == 0 || wl->block == 0 || (wl->block & 63) || wl->block > 256) {
		set_err("bad workload");
		return -1;
	}
	int layout = h->cfg.layout == KMA_LAYOUT_REFERENCE ? 0 : 1;
	unsigned long long *d_stats = NULL;
	hipStream_t s = NULL;
	hipEvent_t e0 = NULL, e1 = NULL;
	int rc = -1;
	uint64_t st0[NAT_NSTATS], st1[NAT_NSTATS];
	unsigned long long hs[4];
	hipGraph_t graph = NULL;
	hipGraphExec_t gexec = NULL;
	if (read_native_stats(h, st0) != 0)
		return -1;
	if (hipMalloc((void **)&d_stats, 4 * 8) != hipSuccess || hipStreamCreate(&s) != hipSuccess ||
	    hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
		set_err("workload setup failed");
		goto done;
	}
	if (wl->form == KMA_FORM_BATCHED && layout != 1) {
		set_err("the batched form needs a native-layout heap");
		goto done;
	}
	/* warm-up rounds, untimed, not counted */
	if (hipMemset(d_stats, 0, 32) != hipSuccess)
		goto done;
	if (wl->warmup) {
		if (wl->form == KMA_FORM_BATCHED)
			rc = kmad_batched(h->dev, wl->round0, wl->warmup, (int)wl->mode, wl->fixed, wl->seed, wl->nwi,
					  d_stats, s);
		else
			rc = kmad_rounds(layout, h->dev, wl->round0, wl->warmup, (int)wl->mode, wl->fixed, wl->seed,
					 wl->nwi, wl->block, d_stats, s);
		if (rc) {
			set_err("warm-up launch failed: %d", rc);
			goto done;
		}
	}
	if (hipStreamSynchronize(s) != hipSuccess || hipMemset(d_stats, 0, 32) != hipSuccess) {
		set_err("warm-up failed");
		rc = -1;
		goto done;
	}
	if (read_native_stats(h, st0) != 0) {
		rc = -1;
		goto done;
	}
	uint32_t r0 = wl->round0 + wl->warmup;
	if (wl->form == KMA_FORM_BATCHED) {
		/* one persistent launch for all rounds, captured in a graph */
		if (hipStreamBeginCapture(s, hipStreamCaptureModeGlobal) != hipSuccess) {
			set_err("graph capture failed");
			rc = -1;
			goto done;
		}
		rc = kmad_batched(h->dev, r0, wl->rounds, (int)wl->mode, wl->fixed, wl->seed, wl->nwi, d_stats, s);
		if (hipStreamEndCapture(s, &graph) != hipSuccess || rc != 0 ||
		    hipGraphInstantiate(&gexec, graph, NULL, NULL, 0) != hipSuccess) {
			set_err("graph build failed: %d", rc);
			rc = -1;
			goto done;
		}
	}
	hipEventRecord(e0, s);
	if (wl->form == KMA_FORM_PLAIN) {
		for (uint32_t r = 0; r < wl->rounds; r++) {
			rc = kmad_rounds(layout, h->dev, r0 + r, 1, (int)wl->mode, wl->fixed, wl->seed, wl->nwi,
					 wl->block, d_stats, s);
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
		rc = kmad_rounds(layout, h->dev, r0, wl->rounds, (int)wl->mode, wl->fixed, wl->seed, wl->nwi,
				 wl->block, d_stats, s);
		if (rc) {
			set_err("launch failed: %d", rc);
			goto done;
		}
		out->launches = 1;
	}
	hipEventRecord(e1, s);
	if (hipEventSynchronize(e1) != hipSuccess) {
		set_err("workload kernels failed: %s", hipGetErrorString(hipGetLastError()));
		rc = -1;
		goto done;
	}
	float ms = 0;
	hipEventElapsedTime(&ms, e0, e1);
	out->gpu_ms = ms;
	out->kernel_ms_avg = out->launches ? ms / (double)out->launches : 0;
	if (hipMemcpy(hs, d_stats, 32, hipMemcpyDeviceToHost) != hipSuccess || read_native_stats(h, st1) != 0) {
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
	if (s)
		hipStreamDestroy(s);
	if (d_stats)
		hipFree(d_stats);
	return rc;
}

int kma_atomic_peak(uint64_t bytes, uint64_t nthreads, uint32_t iters, int pattern, uint32_t warmup,
		    uint32_t reps, double *ops_per_s, double *ms_out)
{
	if (bytes < 4096 || nthreads == 0 || iters == 0 || reps == 0 || pattern < 0 || pattern > 2) {
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
