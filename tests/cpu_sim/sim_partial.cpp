/*
 * CPU simulation of partial-superblock reuse on the native layout.
 *
 * Every lane keeps up to `ring` blocks of one size class with random
 * lifetimes (1..hold_max rounds): each round it frees the blocks whose
 * lifetime ended, then allocates one block.  Blocks of one superblock
 * therefore die at different times, so superblocks detached from their
 * heads drain part-way long before they empty; with the heap sized so that
 * the live blocks fill about 70 % of it, refills must attach such partly
 * free superblocks again (the partial index) to keep allocations
 * succeeding.
 *
 * usage: kma_sim_partial sb_log2 sblocks nwg rounds seed nctx class_size hold_max ring variant
 *   variant 0: per-lane kma_malloc / kma_free; 1: kma_wg_malloc / kma_wg_free
 *   with work-group staging (16-lane claim groups)
 *   KMA_SIM_DUMP=path writes the metadata image before the final frees to
 *   path.run (live blocks and listed superblocks present), the image of the
 *   fresh heap to path.init and the image after the final frees and the
 *   trim to path.final
 * prints: mallocs nulls bad peak_live capacity listed reattached canonical_problems
 */
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "kma_layout.h"
#include "kma_workloads.h"

thread_local sim_lane g_sim_lane;
dim3 g_sim_block_dim, g_sim_grid_dim;

namespace {

constexpr uint32_t BLOCK = KMA_BATCH_BLOCK;
constexpr uint32_t MAX_RING = 64;

struct held {
	uint32_t *p;
	uint32_t expire;
	uint32_t pat;
};

uint64_t mix(uint64_t a, uint64_t b, uint64_t c)
{
	uint64_t z = a * 0x9E3779B97F4A7C15ull + b * 0xBF58476D1CE4E5B9ull + c * 0x94D049BB133111EBull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

int canonical_problems(kma_dheap *H)
{
	using namespace kma_nat;
	int probs = 0;
	for (uint64_t sb = 0; sb < H->nsb; sb++) {
		if (*state(H, sb) & ((1ull << 42) - 1))
			probs++;
		for (uint32_t w = 0; w < H->bm_words; w++)
			probs += bitmap(H, sb)[w] != 0;
	}
	for (uint32_t x = 0; x < KMA_NXCD; x++)
		for (uint32_t c = 0; c < H->nclasses; c++)
			for (uint32_t i = 0; i < H->pidx_words; i++)
				probs += pidx(H, x, c)[i] != 0;
	std::vector<uint8_t> seen(H->nsb, 0);
	uint64_t n = 0;
	for (uint32_t s = 0; s < H->npool; s++) {
		uint32_t top = (uint32_t)*pool(H, s);
		while (top && n <= H->nsb) {
			probs += seen[top - 1]++ != 0;
			n++;
			top = *pnext(H, top - 1);
		}
	}
	probs += n != H->nsb;
	return probs;
}

void trim(kma_dheap *H)
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
}

bool dump_image(const char *dump, const char *suffix, const char *buf, uint64_t n)
{
	std::string name = std::string(dump) + suffix;
	FILE *f = fopen(name.c_str(), "wb");
	bool ok = f && fwrite(buf, 1, n, f) == n;
	if (f)
		ok = fclose(f) == 0 && ok;
	return ok;
}

} // namespace

int main(int argc, char **argv)
{
	if (argc < 11) {
		fprintf(stderr, "usage: %s sb_log2 sblocks nwg rounds seed nctx class_size hold_max ring variant\n", argv[0]);
		return 2;
	}
	kma_config in;
	memset(&in, 0, sizeof(in));
	in.layout = KMA_LAYOUT_NATIVE;
	in.sb_log2 = (uint32_t)atoi(argv[1]);
	in.sblocks = strtoull(argv[2], 0, 0);
	const uint32_t nwg = (uint32_t)atoi(argv[3]), rounds = (uint32_t)atoi(argv[4]);
	const uint64_t seed = strtoull(argv[5], 0, 0);
	in.nctx = (uint32_t)atoi(argv[6]);
	in.nclasses = 1;
	in.class_size[0] = (uint32_t)atoi(argv[7]);
	const uint32_t hold_max = (uint32_t)atoi(argv[8]), ring = (uint32_t)atoi(argv[9]);
	const int variant = atoi(argv[10]);
	if (ring == 0 || ring > MAX_RING || hold_max == 0) {
		fprintf(stderr, "ring 1..%u, hold_max >= 1\n", MAX_RING);
		return 2;
	}
	const char *dump = getenv("KMA_SIM_DUMP");
	kma_config cfg;
	nat_hdr nh;
	char err[200];
	if (kma_native_layout(&in, &cfg, &nh, err, sizeof(err)) != 0) {
		fprintf(stderr, "layout: %s\n", err);
		return 2;
	}
	const uint64_t total = kma_native_alloc_bytes(&nh);
	char *buf = (char *)aligned_alloc(4096, (total + 4095) & ~4095ull);
	memset(buf, 0, total);
	kma_native_place(&nh, (uint64_t)(uintptr_t)buf, total);
	memcpy(buf, &nh, sizeof(nh));
	kma_dheap *H = (kma_dheap *)buf;
	{
		const uint64_t n = kma_nat::init_items(H->nsb, H->npool, H->npages);
		for (uint64_t i = 0; i < n; i++)
			kma_nat::init_item(H, i);
	}
	if (dump && !dump_image(dump, ".init", buf, nh.off_data)) {
		fprintf(stderr, "cannot write the image\n");
		return 3;
	}
	const uint32_t size = H->class_size[0];
	const uint64_t capacity = H->nsb * H->class_cap[0];

	g_sim_block_dim = {BLOCK, 1, 1};
	g_sim_grid_dim = {nwg, 1, 1};
	std::atomic<uint64_t> mallocs{0}, nulls{0}, bad{0}, live{0}, peak{0};
	std::vector<std::unique_ptr<sim_wg>> wgs;
	std::vector<std::unique_ptr<sim_wave>> waves;
	std::vector<std::unique_ptr<kma_wg_lds<BLOCK>>> lds;
	std::unique_ptr<sim_wg> all(new sim_wg(nwg * BLOCK)); /* the dump point: every lane */
	for (uint32_t b = 0; b < nwg; b++) {
		wgs.emplace_back(new sim_wg(BLOCK));
		lds.emplace_back(new kma_wg_lds<BLOCK>());
		for (uint32_t w = 0; w < BLOCK / 64; w++)
			waves.emplace_back(new sim_wave());
	}
	bool dumped = true;
	std::vector<std::thread> th;
	for (uint32_t b = 0; b < nwg; b++)
		for (uint32_t t = 0; t < BLOCK; t++)
			th.emplace_back([&, b, t]() {
				g_sim_lane.tid = {t, 0, 0};
				g_sim_lane.bid = {b, 0, 0};
				g_sim_lane.wg = wgs[b].get();
				g_sim_lane.wave = waves[b * (BLOCK / 64) + t / 64].get();
				const uint64_t gid = (uint64_t)b * BLOCK + t;
				kma_wg_lds<BLOCK> &L = *lds[b];
				held h[MAX_RING];
				memset(h, 0, sizeof(h));
				auto free_one = [&](uint32_t *p) {
					if (variant == 1)
						kma_wg_free<BLOCK, 16u>(H, p, L);
					else
						kma_free(H, p);
				};
				for (uint32_t r = 0; r < rounds; r++) {
					for (uint32_t k = 0; k < ring; k++) {
						uint32_t *p = h[k].p && h[k].expire <= r ? h[k].p : nullptr;
						if (p) {
							bad += p[0] != h[k].pat || p[size / 4 - 1] != ~h[k].pat;
							h[k].p = nullptr;
							live--;
						}
						free_one(p);
					}
					uint32_t k = 0;
					while (k < ring && h[k].p)
						k++;
					const uint32_t want = k < ring ? size : 0u;
					uint32_t *p = (uint32_t *)(variant == 1 ? kma_wg_malloc<BLOCK, 16u>(H, want, L)
										: kma_malloc(H, want));
					if (want) {
						mallocs++;
						if (!p) {
							nulls++;
						} else {
							const uint32_t pat = (uint32_t)mix(seed, gid, r);
							p[0] = pat;
							p[size / 4 - 1] = ~pat;
							h[k] = {p, r + 1u + (uint32_t)(mix(seed, gid, r + 0x9999u) % hold_max), pat};
							const uint64_t lv = ++live;
							uint64_t pk = peak.load();
							while (lv > pk && !peak.compare_exchange_weak(pk, lv)) {
							}
						}
					}
				}
				/* quiescent point: every lane has finished its rounds */
				all->bar.arrive_and_wait();
				if (gid == 0 && dump) {
					std::string name = std::string(dump) + ".run";
					FILE *f = fopen(name.c_str(), "wb");
					dumped = f && fwrite(buf, 1, nh.off_data, f) == nh.off_data;
					if (f)
						dumped = fclose(f) == 0 && dumped;
				}
				all->bar.arrive_and_wait();
				for (uint32_t k = 0; k < ring; k++) {
					uint32_t *p = h[k].p;
					if (p) {
						bad += p[0] != h[k].pat || p[size / 4 - 1] != ~h[k].pat;
						live--;
					}
					free_one(p);
				}
			});
	for (auto &t : th)
		t.join();
	if (!dumped) {
		fprintf(stderr, "cannot write the image\n");
		return 3;
	}
	const uint64_t *st = kma_nat::stats(H);
	const uint64_t listed = st[KMA_STAT_LISTED], reattached = st[KMA_STAT_REATTACHED];
	trim(H);
	const int probs = canonical_problems(H);
	if (dump && !dump_image(dump, ".final", buf, nh.off_data)) {
		fprintf(stderr, "cannot write the image\n");
		return 3;
	}
	printf("mallocs %llu nulls %llu bad %llu peak_live %llu capacity %llu listed %llu reattached %llu "
	       "double_frees %llu waits %llu canonical_problems %d\n",
	       (unsigned long long)mallocs.load(), (unsigned long long)nulls.load(), (unsigned long long)bad.load(),
	       (unsigned long long)peak.load(), (unsigned long long)capacity, (unsigned long long)listed,
	       (unsigned long long)reattached, (unsigned long long)st[KMA_STAT_DOUBLE_FREE],
	       (unsigned long long)(st[KMA_STAT_WAIT_RESERVE] + st[KMA_STAT_WAIT_CLAIM] + st[KMA_STAT_WAIT_WGCLAIM] +
				    st[KMA_STAT_WAIT_POOL]),
	       probs);
	free(buf);
	return 0;
}
