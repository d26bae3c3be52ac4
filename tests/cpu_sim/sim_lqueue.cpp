/*
 * CPU simulation of the linked pointer queue (include/kma_lqueue.h): the
 * device header built against the CPU stand-in for hip_runtime.h, one OS
 * thread per device thread, running the queue test kernels' bodies:
 * every thread enqueues its item, then `rounds` times dequeues one and
 * enqueues it again (all lanes of a wavefront call together; a NULL item is
 * a no-op).  Afterwards the chain from head is walked as the reference's
 * validator does (tb_clQueue.c:126-157).
 *
 * usage: kma_sim_lqueue nwg rounds [nq]
 *   nq queues (default 1): each call names a queue hashed from the thread
 *   and the round, so a wavefront's calls mix queues
 * prints: items, stubs (1: every queue holds its stub once), tail_ok (every
 *   tail names its chain's last node), cycle, outside, repeats, lost
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <memory>
#include <thread>
#include <vector>

#include "kma_lqueue.h"

thread_local sim_lane g_sim_lane;
dim3 g_sim_block_dim, g_sim_grid_dim;

int main(int argc, char **argv)
{
	if (argc < 3) {
		fprintf(stderr, "usage: %s nwg rounds [nq]\n", argv[0]);
		return 2;
	}
	const uint32_t nwg = (uint32_t)atoi(argv[1]), rounds = (uint32_t)atoi(argv[2]), B = 256;
	const uint32_t nq = argc > 3 ? (uint32_t)atoi(argv[3]) : 1u;
	if (nq == 0) {
		fprintf(stderr, "nq >= 1\n");
		return 2;
	}
	const uint64_t n = (uint64_t)nwg * B;
	kma_dlqueue *qs = (kma_dlqueue *)aligned_alloc(128, nq * sizeof(kma_dlqueue));
	memset(qs, 0, nq * sizeof(kma_dlqueue));
	for (uint32_t i = 0; i < nq; i++)
		qs[i].head = qs[i].tail = (uint64_t)(uintptr_t)&qs[i].stub;
	uint64_t *items = (uint64_t *)aligned_alloc(16, n * 16);
	memset(items, 0, n * 16);
	g_sim_block_dim = {B, 1, 1};
	g_sim_grid_dim = {nwg, 1, 1};
	std::atomic<uint64_t> lost{0};
	std::vector<std::unique_ptr<sim_wg>> wgs;
	std::vector<std::unique_ptr<sim_wave>> waves;
	for (uint32_t b = 0; b < nwg; b++) {
		wgs.emplace_back(new sim_wg(B));
		for (uint32_t w = 0; w < B / 64; w++)
			waves.emplace_back(new sim_wave());
	}
	/* with several queues, a lane's queue for each call is a hash of its
	 * id and the round: a wavefront's calls then mix queues */
	auto pick = [nq](uint64_t gid, uint32_t r, uint32_t salt) {
		uint64_t z = (gid + 1) * 0x9E3779B97F4A7C15ull ^ ((uint64_t)r << 32 | salt) * 0xBF58476D1CE4E5B9ull;
		z ^= z >> 29;
		return (uint32_t)(z % nq);
	};
	std::vector<std::thread> th;
	for (uint32_t b = 0; b < nwg; b++)
		for (uint32_t t = 0; t < B; t++)
			th.emplace_back([&, b, t]() {
				g_sim_lane.tid = {t, 0, 0};
				g_sim_lane.bid = {b, 0, 0};
				g_sim_lane.wg = wgs[b].get();
				g_sim_lane.wave = waves[b * (B / 64) + t / 64].get();
				const uint64_t gid = (uint64_t)b * B + t;
				void *mine = items + 2u * gid;
				if (!kma_lq_enqueue(&qs[pick(gid, 0, 1)], mine))
					lost++;
				for (uint32_t r = 0; r < rounds; r++) {
					void *it = kma_lq_dequeue(&qs[pick(gid, r, 2)]);
					if (!kma_lq_enqueue(&qs[pick(gid, r, 3)], it) && it)
						lost++;
				}
			});
	for (auto &x : th)
		x.join();
	/* walk every queue */
	const uint64_t P = (1ull << 48) - 1;
	const uint64_t lo = (uint64_t)(uintptr_t)items, hi = lo + n * 16;
	std::vector<uint8_t> seen(n, 0);
	uint64_t cnt = 0, stubs_bad = 0, outside = 0, repeats = 0, cycles = 0, tails_bad = 0;
	for (uint32_t i = 0; i < nq; i++) {
		kma_dlqueue *q = &qs[i];
		const uint64_t stub = (uint64_t)(uintptr_t)&q->stub;
		uint64_t cur = q->head & P, last = 0, stubs = 0, steps = 0;
		bool out = false;
		while (cur && steps <= n + 1) {
			if (cur == stub) {
				stubs++;
			} else if (cur >= lo && cur < hi && (cur - lo) % 16 == 0) {
				if (seen[(cur - lo) / 16]++)
					repeats++;
				cnt++;
			} else {
				outside++;
				out = true;
				break;
			}
			last = cur;
			cur = *(uint64_t *)(uintptr_t)cur & P;
			steps++;
		}
		const bool cycle = cur != 0 && !out;
		cycles += cycle;
		stubs_bad += stubs != 1;
		tails_bad += cycle || out || last != (q->tail & P);
	}
	printf("items %llu stubs %llu tail_ok %d cycle %llu outside %llu repeats %llu lost %llu\n",
	       (unsigned long long)cnt, (unsigned long long)(stubs_bad ? 0 : 1), tails_bad ? 0 : 1,
	       (unsigned long long)cycles, (unsigned long long)outside, (unsigned long long)repeats,
	       (unsigned long long)lost.load());
	fflush(stdout);
	free(items);
	free(qs);
	return 0;
}
