/*
 * kma_internal.h - definitions shared by the host C files of libkma.so
 * (not part of the public API).
 */
#ifndef KMA_INTERNAL_H
#define KMA_INTERNAL_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "kma.h"
#include "kma_layout.h"

#define NAT_NSTATS 11u
#define NAT_STAT_DOUBLE_FREE 0
#define NAT_STAT_BAD_FREE 1

/* Reference layout: stats trailer after the image (kma_ref_device.h,
 * kma_ref::T_* word indexes): at align64(image bytes), 256 bytes. */
#define REF_TRAILER_BYTES 256u
#define REF_TRAILER_OFF(bytes) (((bytes) + 63u) & ~63ull)
#define REF_T_WORDS 8u

struct kma_heap {
	kma_config cfg;
	char *dev;        /* device buffer == header address */
	uint64_t bytes;   /* bytes allocated */
	nat_hdr nh;       /* native header (host copy) */
};

/* Set the calling thread's error message (kma_last_error). */
__attribute__((visibility("hidden"), format(printf, 1, 2))) void kma_i_set_err(const char *fmt, ...);
#define set_err kma_i_set_err

/* The native heap's 8 counters (zeros for the reference layout). */
__attribute__((visibility("hidden"))) int kma_i_read_native_stats(kma_heap *h, uint64_t *st);

/* The reference heap's trailer words (REF_T_WORDS; zeros for the native
 * layout). */
__attribute__((visibility("hidden"))) int kma_i_read_ref_trailer(kma_heap *h, uint64_t *tw);

/* Nanoseconds per tick of the device's constant wall clock. */
__attribute__((visibility("hidden"))) double kma_i_wall_clock_ns(void);

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

#endif
