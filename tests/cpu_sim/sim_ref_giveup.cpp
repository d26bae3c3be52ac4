/*
 * CPU check of the reference-layout allocator's bounded waits
 * (include/kma_ref_device.h), one simulated work-item, built with a 50 ms
 * per-call watchdog:
 *  deadline  an armed deadline that has passed: malloc returns NULL at
 *            once, counted as a deadline give-up, nothing changes;
 *  claim     a superblock whose bitmap is planted full while its free count
 *            says slots are free: the claim gives up, returns NULL, the
 *            reserved slot goes back into the count (state word unchanged)
 *            and the give-up is counted;
 *  exhaust   an exhausted heap: NULL without taking the class head (it stays
 *            0, no POISON) and without a give-up; after freeing everything
 *            the oracle walks all S-1 superblocks in the free queue.
 * Prints one "name ok" line per case, or "name FAIL ..." and exits 1.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kma_workloads.h"

extern "C" {
#include "refmodel.h"
}

thread_local sim_lane g_sim_lane;
dim3 g_sim_block_dim, g_sim_grid_dim;

static int fails = 0;
#define CHECK(name, cond, ...)                                  \
	do {                                                    \
		if (!(cond)) {                                  \
			printf("%s FAIL: ", name);              \
			printf(__VA_ARGS__);                    \
			printf("\n");                           \
			fails++;                                \
			return;                                 \
		}                                               \
	} while (0)

static const uint32_t LOG2 = 12, S = 8;
static uint8_t *img;
static uint64_t bytes;
static rm_heap h;

static uint64_t *trl() { return (uint64_t *)(img + ((bytes + 63) & ~63ull)); }
static uint32_t u32at(uint64_t off) { uint32_t v; memcpy(&v, img + off, 4); return v; }
static uint64_t u64at(uint64_t off) { uint64_t v; memcpy(&v, img + off, 8); return v; }

static void fresh()
{
	uint64_t alloc = (((bytes + 63) & ~63ull) + 256 + 4095) & ~4095ull;
	memset(img, 0, alloc);
	if (rm_init(&h, LOG2, S, (uint64_t)(uintptr_t)img, img) != RM_OK) {
		printf("rm_init failed\n");
		exit(2);
	}
}

static void case_deadline()
{
	fresh();
	trl()[kma_ref::T_DEADLINE] = 1; /* tick 1: long past */
	void *p = kma_malloc((clheap *)img, 64);
	CHECK("deadline", p == nullptr, "malloc returned %p", p);
	CHECK("deadline", trl()[kma_ref::T_DLHIT] == 1, "deadline hits %llu", (unsigned long long)trl()[kma_ref::T_DLHIT]);
	CHECK("deadline", rm_queue_walk(&h) == S - 1, "queue %lld", (long long)rm_queue_walk(&h));
	trl()[kma_ref::T_DEADLINE] = 0;
	p = kma_malloc((clheap *)img, 64);
	CHECK("deadline", p != nullptr, "malloc after disarming returned NULL");
	kma_free((clheap *)img, p);
	CHECK("deadline", rm_queue_walk(&h) == S - 1, "queue %lld after free", (long long)rm_queue_walk(&h));
	printf("deadline ok\n");
}

static void case_claim()
{
	fresh();
	char *p = (char *)kma_malloc((clheap *)img, 64); /* class 4: 63 slots */
	CHECK("claim", p != nullptr, "first malloc NULL");
	uint64_t hdr = 32 + 8 * (LOG2 - 1);
	uint64_t sboff = hdr + (((uint64_t)(p - (char *)img) - hdr) & ~((1ull << LOG2) - 1));
	uint32_t state0 = u32at(sboff + 4);
	CHECK("claim", (state0 & 0xffffu) == 62, "state %08x", state0);
	/* plant: every slot's bit set (2 bitmap words at the superblock's end) */
	uint32_t *w0 = (uint32_t *)(img + sboff + (1u << LOG2) - 4), *w1 = w0 - 1;
	uint32_t s0 = *w0, s1 = *w1;
	*w0 = 0xffffffffu;
	*w1 = 0x7fffffffu;
	void *q = kma_malloc((clheap *)img, 64);
	CHECK("claim", q == nullptr, "malloc on a full bitmap returned %p", q);
	CHECK("claim", trl()[kma_ref::T_CLAIM] == 1, "claim give-ups %llu", (unsigned long long)trl()[kma_ref::T_CLAIM]);
	CHECK("claim", u32at(sboff + 4) == state0, "state %08x, was %08x: the reserved slot was not returned",
	      u32at(sboff + 4), state0);
	*w0 = s0;
	*w1 = s1;
	kma_free((clheap *)img, p);
	CHECK("claim", rm_queue_walk(&h) == S - 1, "queue %lld after free", (long long)rm_queue_walk(&h));
	CHECK("claim", u64at(32 + 8 * 4) == 0, "class head not detached");
	printf("claim ok\n");
}

static void case_exhaust()
{
	fresh();
	void *blk[64];
	int n = 0;
	for (; n < 64; n++) {
		blk[n] = kma_malloc((clheap *)img, 2040); /* class 9 at 4 KiB: 2 slots */
		if (!blk[n])
			break;
	}
	CHECK("exhaust", n == 2 * (int)(S - 2), "%d blocks before NULL", n);
	CHECK("exhaust", u64at(32 + 8 * 9) == 0, "class head %llx (POISON or attached)",
	      (unsigned long long)u64at(32 + 8 * 9));
	uint64_t g = 0;
	for (int k = 1; k < 8; k++)
		g += trl()[k];
	CHECK("exhaust", g == 0, "%llu give-ups counted for a plain exhaustion", (unsigned long long)g);
	for (int k = 0; k < n; k++)
		kma_free((clheap *)img, blk[k]);
	CHECK("exhaust", rm_queue_walk(&h) == S - 1, "queue %lld after freeing all", (long long)rm_queue_walk(&h));
	printf("exhaust ok\n");
}

int main()
{
	bytes = rm_heap_bytes(LOG2, S);
	img = (uint8_t *)aligned_alloc(4096, (((bytes + 63) & ~63ull) + 256 + 4095) & ~4095ull);
	g_sim_block_dim = {64, 1, 1};
	g_sim_grid_dim = {1, 1, 1};
	g_sim_lane.tid = {0, 0, 0};
	g_sim_lane.bid = {0, 0, 0};
	case_deadline();
	case_claim();
	case_exhaust();
	rm_release(&h);
	free(img);
	return fails ? 1 : 0;
}
