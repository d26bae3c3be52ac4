/*
 * Brute-force check of the bit helpers of the native allocator
 * (include/kma_device.h), built for the CPU like the simulation: nth_lane
 * (lane of the k-th set bit of a 64-bit mask), nth_bit (the same for a
 * 32-bit word) and range_bits (the bits of [lo, hi) inside bitmap word wi),
 * each against a plain loop over every bit, on exhaustive small cases and
 * random masks.  Prints "helpers ok" and exits 0, or names the first
 * mismatch and exits 1.
 */
#include <stdio.h>
#include <stdint.h>

#include "kma_device.h"

thread_local sim_lane g_sim_lane;
dim3 g_sim_block_dim, g_sim_grid_dim;

static uint32_t ref_nth64(uint64_t m, uint32_t k)
{
	for (uint32_t i = 0; i < 64; i++)
		if ((m >> i) & 1u) {
			if (k == 0)
				return i;
			k--;
		}
	return 64;
}

static uint32_t ref_range(uint32_t wi, uint32_t lo, uint32_t hi)
{
	uint32_t r = 0;
	for (uint32_t b = 0; b < 32; b++) {
		uint32_t s = wi * 32 + b;
		if (s >= lo && s < hi)
			r |= 1u << b;
	}
	return r;
}

int main()
{
	using namespace kma_nat;
	uint64_t x = 0x9E3779B97F4A7C15ull;
	for (int t = 0; t < 200000; t++) {
		x ^= x << 13;
		x ^= x >> 7;
		x ^= x << 17;
		uint64_t m = t < 64 ? (~0ull >> t) : (t < 128 ? (1ull << (t - 64)) : x);
		uint32_t n = (uint32_t)__builtin_popcountll(m);
		for (uint32_t k = 0; k < n; k++) {
			if (nth_lane(m, k) != ref_nth64(m, k)) {
				printf("nth_lane(%#llx, %u) = %u, expected %u\n", (unsigned long long)m, k, nth_lane(m, k),
				       ref_nth64(m, k));
				return 1;
			}
			uint32_t w = (uint32_t)m;
			if (k < (uint32_t)__builtin_popcount(w) && nth_bit(w, k) != ref_nth64(w, k)) {
				printf("nth_bit(%#x, %u) = %u, expected %u\n", w, k, nth_bit(w, k), ref_nth64(w, k));
				return 1;
			}
		}
	}
	for (uint32_t wi = 0; wi < 6; wi++)
		for (uint32_t lo = 0; lo <= 192; lo++)
			for (uint32_t hi = lo; hi <= 192; hi++)
				if (range_bits(wi, lo, hi) != ref_range(wi, lo, hi)) {
					printf("range_bits(%u, %u, %u) = %#x, expected %#x\n", wi, lo, hi,
					       range_bits(wi, lo, hi), ref_range(wi, lo, hi));
					return 1;
				}
	printf("helpers ok\n");
	return 0;
}
