/*
 * CPU stand-in for <hip/hip_runtime.h>, used only to build this library's
 * device headers for the CPU simulation in tests/cpu_sim (no GPU, no HIP).
 *
 * Every device thread is one OS thread.  Wavefront-wide operations
 * (shuffles, readlane, ballot) exchange values through a 64-thread barrier,
 * so they are only valid where all 64 lanes of the wavefront execute them
 * together -- which is how the work-group collective paths use them.
 * __syncthreads() is a barrier over the work-group's threads.  Atomics are
 * sequentially consistent GCC atomics on host memory.
 */
#ifndef KMA_CPU_SIM_HIP_RUNTIME_H
#define KMA_CPU_SIM_HIP_RUNTIME_H

#include <stdint.h>
#include <stdlib.h>

#include <barrier>
#include <chrono>
#include <thread>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)

struct dim3 {
	uint32_t x, y, z;
};

struct sim_wave {
	std::barrier<> bar{64};
	uint64_t slot[64];
};

struct sim_wg {
	explicit sim_wg(uint32_t n) : bar(n) {}
	std::barrier<> bar;
};

struct sim_lane {
	dim3 tid, bid;
	sim_wave *wave;
	sim_wg *wg;
};

extern thread_local sim_lane g_sim_lane;
extern dim3 g_sim_block_dim, g_sim_grid_dim;

#define threadIdx (g_sim_lane.tid)
#define blockIdx (g_sim_lane.bid)
#define blockDim (g_sim_block_dim)
#define gridDim (g_sim_grid_dim)

static inline uint32_t __lane_id() { return g_sim_lane.tid.x & 63u; }

template <typename T>
static inline T sim_exchange(T v, uint32_t src)
{
	sim_wave *w = g_sim_lane.wave;
	uint64_t raw = 0;
	static_assert(sizeof(T) <= 8, "exchange width");
	__builtin_memcpy(&raw, &v, sizeof(T));
	w->slot[__lane_id()] = raw;
	w->bar.arrive_and_wait();
	uint64_t got = w->slot[src & 63u];
	w->bar.arrive_and_wait();
	T r;
	__builtin_memcpy(&r, &got, sizeof(T));
	return r;
}

/* width splits the wavefront into segments of `width` lanes, as on the
 * device: sources are taken within the caller's segment */
template <typename T>
static inline T __shfl(T v, int src, int width = 64)
{
	uint32_t l = __lane_id(), w = (uint32_t)width;
	return sim_exchange(v, (l & ~(w - 1u)) + ((uint32_t)src & (w - 1u)));
}
template <typename T>
static inline T __shfl_up(T v, unsigned int d, int width = 64)
{
	uint32_t l = __lane_id(), w = (uint32_t)width;
	return sim_exchange(v, (l & (w - 1u)) >= d ? l - d : l);
}
template <typename T>
static inline T __shfl_xor(T v, int m, int width = 64)
{
	(void)width;
	return sim_exchange(v, __lane_id() ^ (uint32_t)m);
}
static inline int sim_readlane(int v, int l) { return sim_exchange(v, (uint32_t)l); }
#define __builtin_amdgcn_readlane(v, l) sim_readlane((v), (l))

static inline unsigned long long __ballot(int pred)
{
	sim_wave *w = g_sim_lane.wave;
	w->slot[__lane_id()] = pred ? 1u : 0u;
	w->bar.arrive_and_wait();
	unsigned long long m = 0;
	for (int i = 0; i < 64; i++)
		if (w->slot[i])
			m |= 1ull << i;
	w->bar.arrive_and_wait();
	return m;
}

static inline void __syncthreads() { g_sim_lane.wg->bar.arrive_and_wait(); }

static inline uint64_t wall_clock64()
{
	using namespace std::chrono;
	return (uint64_t)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() / 10u;
}
static inline uint64_t clock64() { return wall_clock64(); }
#define __builtin_amdgcn_s_sleep(n) std::this_thread::yield()
/* fences: sequentially consistent on the host */
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(__ATOMIC_SEQ_CST)
#define __builtin_amdgcn_s_waitcnt(n) __atomic_thread_fence(__ATOMIC_SEQ_CST)
/* hwreg reads: only HW_REG_XCC_ID is used; work-groups are spread over 8
 * simulated XCDs round-robin, as the hardware dispatcher does */
#define __builtin_amdgcn_s_getreg(enc) ((int)(g_sim_lane.bid.x % 8u))

static inline int __popc(uint32_t x) { return __builtin_popcount(x); }
static inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
static inline int __ffs(uint32_t x) { return __builtin_ffs((int)x); }
static inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
static inline int __clz(uint32_t x) { return x ? __builtin_clz(x) : 32; }
static inline int __clzll(unsigned long long x) { return x ? __builtin_clzll(x) : 64; }

static inline uint32_t min(uint32_t a, uint32_t b) { return a < b ? a : b; }
static inline uint64_t min(uint64_t a, uint64_t b) { return a < b ? a : b; }
static inline int min(int a, int b) { return a < b ? a : b; }
static inline uint32_t max(uint32_t a, uint32_t b) { return a > b ? a : b; }
static inline uint64_t max(uint64_t a, uint64_t b) { return a > b ? a : b; }
static inline int max(int a, int b) { return a > b ? a : b; }

#define __HIP_MEMORY_SCOPE_WORKGROUP 2
#define __HIP_MEMORY_SCOPE_AGENT 3

template <typename T>
static inline T __hip_atomic_load(const T *p, int, int) { return __atomic_load_n(p, __ATOMIC_SEQ_CST); }
template <typename T, typename U>
static inline void __hip_atomic_store(T *p, U v, int, int) { __atomic_store_n(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename U>
static inline bool __hip_atomic_compare_exchange_strong(T *p, T *e, U d, int, int, int)
{
	return __atomic_compare_exchange_n(p, e, (T)d, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
}
template <typename T, typename U>
static inline T __hip_atomic_exchange(T *p, U v, int, int) { return __atomic_exchange_n(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename U>
static inline T __hip_atomic_fetch_add(T *p, U v, int, int) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename U>
static inline T __hip_atomic_fetch_sub(T *p, U v, int, int) { return __atomic_fetch_sub(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename U>
static inline T __hip_atomic_fetch_or(T *p, U v, int, int) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename U>
static inline T __hip_atomic_fetch_and(T *p, U v, int, int) { return __atomic_fetch_and(p, (T)v, __ATOMIC_SEQ_CST); }

template <typename T, typename U>
static inline T atomicAdd(T *p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename U>
static inline T atomicOr(T *p, U v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST); }
template <typename T, typename U>
static inline T atomicCAS(T *p, U cmp, U val)
{
	T e = (T)cmp;
	__atomic_compare_exchange_n(p, &e, (T)val, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
	return e;
}

#endif
