/*
 * CPU simulation of the device allocator's work-group collective path.
 *
 * Builds the library's own device headers (kma_device.h, kma_wg_device.h,
 * kma_workloads.h) against the CPU stand-in for hip_runtime.h in this
 * directory and runs the batched workload body on a native heap that lives
 * in host memory, one OS thread per device thread.  Used by the CPU tests to
 * check the allocator's invariants without a GPU.
 *
 * usage: kma_sim sb_log2 sblocks nwg nwi rounds mode seed nctx npool [variant]
 * prints: ops nulls bad double_frees bad_frees waits canonical_problems
 */
#include <stdio.h>
#include <string.h>

#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "kma_layout.h"
#include "kma_workloads.h"

thread_local sim_lane g_sim_lane;
dim3 g_sim_block_dim, g_sim_grid_dim;

static void nat_init(kma_dheap *H)
{
	/* the device init, one item at a time */
	uint64_t n = kma_nat::init_items(H->nsb, H->npool, H->npages);
	for (uint64_t i = 0; i < n; i++)
		kma_nat::init_item(H, i);
}

static void trim(kma_dheap *H)
{
	using namespace kma_nat;
	for (uint32_t ctx = 0; ctx < H->nctx; ctx++)
		for (uint32_t c = 0; c < H->nclasses; c++) {
			uint64_t h = __hip_atomic_exchange(head(H, ctx, c), HEAD_EMPTY, 0, 0);
			if (h <= HEAD_BUSY)
				continue;
			uint32_t sb = (uint32_t)h - 2u;
			uint64_t s = f_and(state(H, sb), ~ST_OWNED) & ~ST_OWNED;
			after_detach(H, sb, s, st_home(s), ctx);
		}
	for (uint32_t x = 0; x < KMA_NXCD; x++)
		for (uint32_t c = 0; c < H->nclasses; c++) {
			for (uint32_t i = 0; i < pidx_w0(H); i++)
				trim_pidx_word(H, x, c, i);
			pidx_rebuild(H, x, c);
		}
	clean_all(H);
	for (uint32_t ctx = 0; ctx < H->nctx; ctx++)
		for (uint32_t k = 0; k < 32; k++)
			if (spare(H, ctx)[k])
				fprintf(stderr, "ctx %u spare slot %u still holds %u\n", ctx, k, spare(H, ctx)[k]);
}

static int canonical_problems(kma_dheap *H)
{
	using namespace kma_nat;
	int probs = 0;
	for (uint64_t sb = 0; sb < H->nsb; sb++) {
		if (*state(H, sb) & ((1ull << 42) - 1)) { /* all but the generation */
			if (probs < 5)
				fprintf(stderr, "sb %llu state %#llx\n", (unsigned long long)sb,
					(unsigned long long)*state(H, sb));
			probs++;
		}
		for (uint32_t w = 0; w < H->bm_words; w++)
			if (bitmap(H, sb)[w]) {
				if (probs < 5)
					fprintf(stderr, "sb %llu word %u = %#x\n", (unsigned long long)sb, w,
						bitmap(H, sb)[w]);
				probs++;
			}
	}
	std::vector<uint8_t> seen(H->nsb, 0);
	uint64_t n = 0;
	for (uint32_t s = 0; s < H->npool; s++) {
		uint32_t top = (uint32_t)*pool(H, s);
		uint64_t walked = 0;
		while (top && walked <= H->nsb) {
			if (seen[top - 1]++)
				probs++;
			n++;
			walked++;
			top = *pnext(H, top - 1);
		}
		if (walked != pool(H, s)[1])
			probs++;
	}
	if (n != H->nsb)
		probs++;
	/* partial indexes empty */
	for (uint32_t x = 0; x < KMA_NXCD; x++)
		for (uint32_t c = 0; c < H->nclasses; c++)
			for (uint32_t i = 0; i < H->pidx_words; i++)
				if (pidx(H, x, c)[i]) {
					fprintf(stderr, "xcd %u class %u partial-index word %u = %#llx\n", x, c, i,
						(unsigned long long)pidx(H, x, c)[i]);
					probs++;
				}
	/* every non-empty shard has its hint bit set */
	uint32_t S = H->npool / KMA_NXCD;
	for (uint32_t s = 0; s < H->npool; s++)
		if ((uint32_t)pool(H, s)[0] != 0 && !(*pmask(H, s / S) & ((uint64_t)1 << (s % S)))) {
			fprintf(stderr, "shard %u non-empty but its hint bit is clear\n", s);
			probs++;
		}
	return probs;
}

static bool dump_image(const char *path, const char *suffix, const char *buf, uint64_t n)
{
	std::string name = std::string(path) + suffix;
	FILE *f = fopen(name.c_str(), "wb");
	bool ok = f && fwrite(buf, 1, n, f) == n;
	if (f)
		ok = fclose(f) == 0 && ok;
	if (!ok)
		fprintf(stderr, "cannot write %s\n", name.c_str());
	return ok;
}

int main(int argc, char **argv)
{
	if (argc < 10) {
		fprintf(stderr, "usage: %s sb_log2 sblocks nwg nwi rounds mode seed nctx npool [variant [npages]]\n"
				"  KMA_SIM_DUMP=path writes the metadata image after the run to path.run\n"
				"  and after the trim to path.trim\n",
			argv[0]);
		return 2;
	}
	kma_config in;
	memset(&in, 0, sizeof(in));
	in.layout = KMA_LAYOUT_NATIVE;
	in.sb_log2 = (uint32_t)atoi(argv[1]);
	in.sblocks = strtoull(argv[2], 0, 0);
	uint32_t nwg = (uint32_t)atoi(argv[3]);
	uint64_t nwi = strtoull(argv[4], 0, 0);
	uint32_t rounds = (uint32_t)atoi(argv[5]);
	int mode = atoi(argv[6]);
	uint64_t seed = strtoull(argv[7], 0, 0);
	in.nctx = (uint32_t)atoi(argv[8]);
	in.npool = (uint32_t)atoi(argv[9]);
	const uint32_t variant = argc > 10 ? (uint32_t)atoi(argv[10]) : 0u;
	in.npages = argc > 11 ? strtoull(argv[11], 0, 0) : 0u;
	const char *dump = getenv("KMA_SIM_DUMP");
	kma_config cfg;
	nat_hdr nh;
	char err[200];
	if (kma_native_layout(&in, &cfg, &nh, err, sizeof(err)) != 0) {
		fprintf(stderr, "layout: %s\n", err);
		return 2;
	}
	uint64_t total = kma_native_alloc_bytes(&nh);
	char *buf = (char *)aligned_alloc(4096, (total + 4095) & ~4095ull);
	memset(buf, 0, total);
	kma_native_place(&nh, (uint64_t)(uintptr_t)buf, total);
	memcpy(buf, &nh, sizeof(nh));
	kma_dheap *H = (kma_dheap *)buf;
	nat_init(H);

	g_sim_block_dim = {KMA_BATCH_BLOCK, 1, 1};
	g_sim_grid_dim = {nwg, 1, 1};
	unsigned long long stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	std::vector<std::unique_ptr<sim_wg>> wgs;
	std::vector<std::unique_ptr<sim_wave>> waves;
	std::vector<std::unique_ptr<kma_wg_lds<KMA_BATCH_BLOCK>>> lds;
	std::vector<std::thread> th;
	g_sim_lane.bid = {0, 0, 0};
	for (uint32_t b = 0; b < nwg; b++) {
		wgs.emplace_back(new sim_wg(KMA_BATCH_BLOCK));
		lds.emplace_back(new kma_wg_lds<KMA_BATCH_BLOCK>());
		for (uint32_t w = 0; w < KMA_BATCH_BLOCK / 64; w++)
			waves.emplace_back(new sim_wave());
	}
	for (uint32_t b = 0; b < nwg; b++)
		for (uint32_t t = 0; t < KMA_BATCH_BLOCK; t++)
			th.emplace_back([=, &wgs, &waves, &lds, &stats]() {
				g_sim_lane.tid = {t, 0, 0};
				g_sim_lane.bid = {b, 0, 0};
				g_sim_lane.wg = wgs[b].get();
				g_sim_lane.wave = waves[b * (KMA_BATCH_BLOCK / 64) + t / 64].get();
				kma_wl_args wa = {0, rounds, mode, 64, seed, nwi, variant, stats, nullptr};
				switch ((variant & KMA_VAR_AGG_MASK) >> KMA_VAR_AGG_SHIFT) {
				case 1:
					wl_batched<true, 8u>(H, *lds[b], wa);
					break;
				case 2:
					wl_batched<true, 16u>(H, *lds[b], wa);
					break;
				case 3:
					wl_batched<true, 32u>(H, *lds[b], wa);
					break;
				default:
					wl_batched<true, KMA_AGG_WAVE>(H, *lds[b], wa);
				}
			});
	for (auto &t : th)
		t.join();
	uint64_t *st = kma_nat::stats(H);
	if (dump && !dump_image(dump, ".run", buf, nh.off_data))
		return 3;
	trim(H);
	if (dump && !dump_image(dump, ".trim", buf, nh.off_data))
		return 3;
	int probs = canonical_problems(H);
	printf("ops %llu nulls %llu bad %llu double_frees %llu bad_frees %llu waits %llu %llu %llu %llu canonical_problems %d\n",
	       stats[0], stats[1], stats[2], (unsigned long long)st[0], (unsigned long long)st[1],
	       (unsigned long long)st[3], (unsigned long long)st[4], (unsigned long long)st[5],
	       (unsigned long long)st[6], probs);
	free(buf);
	return 0;
}
