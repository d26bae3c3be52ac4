"""The allocation-free deterministic batch entry point
(kma_batch_malloc_det_noalloc) against the blocking one
(kma_batch_malloc_det): on two heaps driven through the same requests, the
offsets, the NULL count and every byte of the heap image must be equal,
for two calls in a row on the shared static scratch (the second with a
smaller batch, so stale values from the first call would show) and for a
third call that must be refused for its 1-slot-class requests.  Device
buffers for its inputs and outputs are the caller's: here from hipMalloc
through ctypes, outside the library."""
import ctypes

import numpy as np
import pytest

import kma_amd

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Hip:
    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p.value)
        return p.value

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        d = self.alloc(arr.nbytes)
        assert self.L.hipMemcpy(d, arr.ctypes.data, arr.nbytes, _H2D) == 0
        return d

    def get(self, d, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data, d, out.nbytes, _D2H) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _image(heap, log2):
    """Heap image with the header's absolute pointers (free.base and the
    class heads, kma.h:81-85) turned into heap offsets, so that images of
    two heaps at different device addresses compare byte for byte."""
    img = heap.snapshot()
    hdr = img[:32 + 8 * (log2 - 1)].view(np.uint64).copy()
    base = np.uint64(heap.device_ptr)
    for k in [1] + list(range(4, len(hdr))):
        if hdr[k]:
            hdr[k] -= base
    img[:hdr.nbytes] = hdr.view(np.uint8)
    return img


def _noalloc(hip, heap, sizes, stream):
    n = len(sizes)
    d_sizes = hip.put(sizes.astype(np.uint32))
    d_out = hip.alloc(8 * n)
    d_status = hip.alloc(16)
    heap.batch_malloc_det_noalloc(d_sizes, n, d_out, d_status, stream)
    assert hip.L.hipStreamSynchronize(stream) == 0
    return hip.get(d_out, n, np.uint64), hip.get(d_status, 2, np.uint64)


def test_noalloc_batch_equals_blocking_batch():
    hip = _Hip()
    stream = ctypes.c_void_p()
    assert hip.L.hipStreamCreate(ctypes.byref(stream)) == 0
    stream = stream.value
    try:
        # the oversubscribed heap shape: 64 MiB of 16 KiB superblocks
        with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 14, 4096) as a, \
                kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 14, 4096) as b:
            for n, seed in ((300000, 2), (5000, 9)):
                print("noalloc det batch n=%d seed=%d" % (n, seed), flush=True)
                sizes = kma_amd.loguniform_sizes(n, 6, 12, seed)
                out_a, nulls_a, _ = a.batch_malloc_det(sizes)
                out_b, status = _noalloc(hip, b, sizes, stream)
                assert np.array_equal(out_a, out_b), np.nonzero(out_a != out_b)[0][:10]
                assert int(status[0]) == nulls_a and int(status[1]) == 0
                assert (n == 5000) or nulls_a > 0  # the large batch oversubscribes the heap
                assert np.array_equal(_image(a, 14), _image(b, 14))
                a.batch_free_det(out_a)
                b.batch_free_det(out_b)
                assert np.array_equal(_image(a, 14), _image(b, 14))
            # 16368 B is a 1-slot class at 16 KiB: refused through the status word
            print("noalloc det batch refusal", flush=True)
            sizes = np.array([64, 16368, 100, 16368, 16368], dtype=np.uint32)
            with pytest.raises(kma_amd.KmaError):
                a.batch_malloc_det(sizes)
            _, status = _noalloc(hip, b, sizes, stream)
            assert int(status[1]) == 3
    finally:
        hip.L.hipStreamDestroy(stream)
        hip.close()


def test_noalloc_batch_rejects_sizes_and_layouts():
    hip = _Hip()
    try:
        d = hip.alloc(64)
        with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 12, 16) as h:
            for n in (0, kma_amd.DET_NOALLOC_MAX + 1):
                with pytest.raises(kma_amd.KmaError, match="outside"):
                    h.batch_malloc_det_noalloc(d, n, d, d, None)
        with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 64) as h:
            with pytest.raises(kma_amd.KmaError):
                h.batch_malloc_det_noalloc(d, 1, d, d, None)
        assert 0 < kma_amd.lib().kma_det_noalloc_scratch_bytes() < 16 << 20
    finally:
        hip.close()
