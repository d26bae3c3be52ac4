/*
 * refmodel.h - CPU restatement of RSpliet/KMA's superblock allocator.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: tests/, the smoke check and
 * the bench's CPU-baseline leg use it as the checker.  The product (libkma.so)
 * never links, loads or calls it.
 *
 * Two halves:
 *  - a serialized model ("rm_*") that replays one work-item at a time and
 *    reproduces the reference heap image byte for byte (64-bit device layout,
 *    zero-initialised buffer), and
 *  - a multi-threaded restatement ("rmt_*") of the same lock-free protocol with
 *    C11/GCC atomics, used only to time the reference algorithm on host cores.
 *
 * Reference files followed: kma.h:27-29,71-85; kma.c:31-53; kma.cl:27-373;
 * clIndexedQueue.h:38-51; clIndexedQueue.cl:44-209.
 */
#ifndef KMA_REFMODEL_H
#define KMA_REFMODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned through rm_malloc's *status. */
#define RM_OK     0
#define RM_HANG   1  /* reference would spin forever (1-slot class, kma.cl:203-204) */
#define RM_EINVAL 2

typedef struct rm_heap {
	uint8_t *img;       /* heap image exactly as the device buffer holds it */
	uint64_t bytes;     /* heap->bytes == sblocks*SB + sizeof(struct clheap) */
	uint64_t base;      /* device address of img[0]; pointer fields hold base+offset */
	uint32_t log2;      /* superblock size log2 (kma.h:28) */
	uint32_t sb;        /* superblock size (kma.h:27) */
	uint32_t buckets;   /* log2 - 1 (kma.h:29) */
	uint32_t hdr;       /* sizeof(struct clheap) on a 64-bit device (kma.h:81-85) */
	uint32_t nsb;       /* sblocks */
	int owns_img;
} rm_heap;

/* Size-class helpers (kma.cl:55-148). */
uint32_t rm_size_by_sbid(uint32_t log2, int sbid);
int      rm_slots_by_size(uint32_t log2, uint32_t size);
/* Returns -1 for size > SB-16 (kma.cl:126-127).  Sizes 0..3 leave the
 * reference's sbid uninitialised (kma.cl:129-135); this model defines
 * 0 -> -1 (NULL) and 1..3 -> class 0. */
int      rm_sbid_by_size(uint32_t log2, uint64_t size);
uint32_t rm_header_bytes(uint32_t log2);

/* Create + clheap_init (kma.c:31-53, kma.cl:27-52).  img may be NULL (the
 * model allocates a zeroed image) or a caller-provided zeroed buffer of
 * rm_heap_bytes() bytes.  Returns 0 on success. */
uint64_t rm_heap_bytes(uint32_t log2, uint32_t sblocks);
int  rm_init(rm_heap *h, uint32_t log2, uint32_t sblocks, uint64_t base, uint8_t *img);
void rm_release(rm_heap *h);

/* Serialized malloc/free (kma.cl:279-373).  Addresses are device addresses
 * (base + offset); 0 is NULL. */
uint64_t rm_malloc(rm_heap *h, uint64_t size, int *status);
void     rm_free(rm_heap *h, uint64_t addr);

/* Free-list walk in the manner of tb_clIndexedQueue.c:172-210: returns the
 * number of nodes reachable from head (including the dummy), -1 on a cycle,
 * -2 if the last node is not the tail. */
int64_t rm_queue_walk(const rm_heap *h);

/* Replays an op list one work-item at a time.  ops[i] >= 0 is malloc(ops[i])
 * and stores the result in out[i]; ops[i] < 0 is free(out[-ops[i]-1]).
 * Optionally writes the 32-bit value pat[i] into the first word of each
 * successful allocation (user workload write).  Returns number of NULLs, or
 * -1 if the reference would hang. */
int64_t rm_replay(rm_heap *h, const int64_t *ops, uint64_t n, uint64_t *out,
		  const uint32_t *pat);

/* Deterministic per-work-item trace values shared with the device code:
 * splitmix64 over (seed, gid, round, lane-of-draw). */
uint64_t rm_mix(uint64_t seed, uint64_t gid, uint64_t round, uint64_t k);

/* ---- multi-threaded restatement (CPU baseline) ---- */
typedef struct rmt_result {
	double seconds;
	uint64_t ops;       /* mallocs + frees that returned / completed */
	uint64_t nulls;
	uint64_t bad;       /* pattern read-back mismatches */
} rmt_result;

/* Full-occupancy workload on host threads: work-items [0, nwi) are split into
 * contiguous slices, one per thread; each work-item runs `rounds` rounds of
 * malloc(2^e), hold (k 32-bit writes), read back, free.  mode 0 = fixed size
 * `fixed` with one write (plain small workload), mode 1 = log-uniform
 * 8 B..2 KiB with 0..64 holds (full-occupancy workload). */
int rmt_run(uint32_t log2, uint32_t sblocks, uint32_t nthreads, uint64_t nwi,
	    uint32_t rounds, uint32_t round0, int mode, uint32_t fixed,
	    uint64_t seed, rmt_result *res);

#ifdef __cplusplus
}
#endif
#endif
