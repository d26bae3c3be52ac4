/*
 * CPU simulation of the reference-layout device allocator
 * (include/kma_ref_device.h) under contention: one OS thread per device
 * thread runs the per-work-item workload body (kma_workloads.h, the
 * reference's kma_test_malloc / _lowvar / _highvar, kma.cl:424-526) on a
 * reference-layout heap whose initial image the oracle writes (rm_init,
 * the image the GPU's k_ref_init is tested to produce).  The allocator path
 * has no wavefront operations, so work-items that get NULL and stop do not
 * disturb the simulation.  Built twice: plain, and with the host
 * AddressSanitizer, so any access outside the heap buffer is reported.
 *
 * usage: kma_sim_ref log2 sblocks nwg nwi rounds mode
 * prints: ops N nulls N bad N queue N  (queue: free-list length walked
 * afterwards by the oracle, -1 if the queue is corrupt)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <thread>
#include <vector>

#include "kma_workloads.h"

extern "C" {
#include "refmodel.h"
}

thread_local sim_lane g_sim_lane;
dim3 g_sim_block_dim, g_sim_grid_dim;

int main(int argc, char **argv)
{
	if (argc < 7) {
		fprintf(stderr, "usage: %s log2 sblocks nwg nwi rounds mode\n", argv[0]);
		return 2;
	}
	const uint32_t log2 = (uint32_t)atoi(argv[1]), S = (uint32_t)atoi(argv[2]);
	const uint32_t nwg = (uint32_t)atoi(argv[3]);
	const uint64_t nwi = strtoull(argv[4], 0, 0);
	const uint32_t rounds = (uint32_t)atoi(argv[5]);
	const int mode = atoi(argv[6]);
	const uint32_t B = 256;
	uint64_t bytes = rm_heap_bytes(log2, S);
	/* the image, then the 256-byte stats trailer at the next 64-byte boundary */
	uint64_t alloc = (((bytes + 63) & ~63ull) + 256 + 4095) & ~4095ull;
	uint8_t *img = (uint8_t *)aligned_alloc(4096, alloc);
	memset(img, 0, alloc);
	rm_heap h;
	if (rm_init(&h, log2, S, (uint64_t)(uintptr_t)img, img) != RM_OK) {
		fprintf(stderr, "rm_init failed\n");
		return 2;
	}
	g_sim_block_dim = {B, 1, 1};
	g_sim_grid_dim = {nwg, 1, 1};
	unsigned long long stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	std::vector<std::unique_ptr<sim_wg>> wgs;
	std::vector<std::unique_ptr<sim_wave>> waves;
	for (uint32_t b = 0; b < nwg; b++) {
		wgs.emplace_back(new sim_wg(B));
		for (uint32_t w = 0; w < B / 64; w++)
			waves.emplace_back(new sim_wave());
	}
	std::vector<std::thread> th;
	for (uint32_t b = 0; b < nwg; b++)
		for (uint32_t t = 0; t < B; t++)
			th.emplace_back([=, &wgs, &waves, &stats]() {
				g_sim_lane.tid = {t, 0, 0};
				g_sim_lane.bid = {b, 0, 0};
				g_sim_lane.wg = wgs[b].get();
				g_sim_lane.wave = waves[b * (B / 64) + t / 64].get();
				kma_wl_args wa = {0, rounds, mode, 64, 1, nwi, 0, stats, nullptr};
				wl_rounds<clheap, false>((clheap *)img, wa);
			});
	for (auto &t : th)
		t.join();
	int64_t q = rm_queue_walk(&h);
	const uint64_t *trl = (const uint64_t *)(img + ((bytes + 63) & ~63ull));
	printf("ops %llu nulls %llu bad %llu queue %lld giveups %llu\n", stats[0], stats[1], stats[2], (long long)q,
	       (unsigned long long)(trl[1] + trl[2] + trl[3] + trl[4] + trl[5]));
	rm_release(&h);
	free(img);
	return 0;
}
