"""Deterministic batch mode on the GPU (kma_batch_malloc_det /
kma_batch_free_det): one parallel batch must equal the oracle's serialized
replay of the same requests in index order, address for address and byte
for byte of the heap image, NULLs included (the oversubscribed workload:
64 MiB heap of 16 KiB superblocks, 524,288 requests log-uniform over
64 B .. 4 KiB, seed 2)."""
import numpy as np
import pytest

import kma_amd
from oracle import refmodel as rm

pytestmark = pytest.mark.gpu


def _oracle_cycle(o, sizes, pat=None):
    """Serialized mallocs of `sizes` in index order, then the frees in index
    order; returns (malloc offsets, image after the mallocs, image after the
    frees)."""
    n = len(sizes)
    out = np.zeros(n, dtype=np.uint64)
    for i, s in enumerate(sizes):
        a = o.malloc(int(s))
        out[i] = a - o.base if a else 0
    img1 = o.image().copy()
    for i in range(n):
        if out[i]:
            o.free(int(out[i]) + o.base)
    return out, img1, o.image().copy()


def _no_overlap(offs, sizes):
    ok = offs != 0
    o = offs[ok].astype(np.int64)
    s = sizes[ok].astype(np.int64)
    order = np.argsort(o)
    o, s = o[order], s[order]
    return bool((o[1:] >= o[:-1] + s[:-1]).all())


@pytest.mark.parametrize("log2,S,n,lo,hi,seed", [
    (14, 64, 4096, 6, 12, 2),     # oversubscribed, 16 KiB superblocks
    (12, 48, 3000, 2, 10, 7),     # 4 KiB superblocks, sizes 4 B .. 1023 B (no 1-slot class)
])
def test_det_batch_equals_serialized_oracle(log2, S, n, lo, hi, seed):
    print("det batch log2=%d S=%d n=%d" % (log2, S, n), flush=True)
    sizes = kma_amd.loguniform_sizes(n, lo, hi, seed)
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, log2, S) as h:
        o = rm.RefHeap(log2, S, base=h.device_ptr)
        for cycle in range(2):  # the second batch starts from the rotated free queue
            out, nulls, _ = h.batch_malloc_det(sizes)
            oout, img1, img2 = _oracle_cycle(o, sizes)
            assert np.array_equal(out, oout), np.nonzero(out != oout)[0][:10]
            assert nulls == int((oout == 0).sum())
            assert _no_overlap(out, sizes)
            d = np.nonzero(h.snapshot() != img1)[0]
            assert d.size == 0, "cycle %d: image after mallocs differs at %s" % (cycle, d[:16])
            h.batch_free_det(out)
            d = np.nonzero(h.snapshot() != img2)[0]
            assert d.size == 0, "cycle %d: image after frees differs at %s" % (cycle, d[:16])
            sizes = sizes[::-1].copy()


def test_oversubscribed_workload_exact():
    """The full oversubscribed workload: success/NULL counts and both images
    equal the serialized oracle."""
    print("det batch oversubscribed 524288", flush=True)
    log2, S, n = 14, 4096, 524288
    sizes = kma_amd.loguniform_sizes(n, 6, 12, 2)
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, log2, S) as h:
        out, nulls, ms = h.batch_malloc_det(sizes)
        o = rm.RefHeap(log2, S, base=h.device_ptr)
        oout, img1, img2 = _oracle_cycle(o, sizes)
        assert nulls == int((oout == 0).sum()) and nulls > n // 2
        assert np.array_equal(out, oout)
        assert _no_overlap(out, sizes)
        assert np.array_equal(h.snapshot(), img1)
        h.batch_free_det(out)
        assert np.array_equal(h.snapshot(), img2)


def test_det_batch_rejects_one_slot_class_and_native_heap():
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 12, 16) as h:
        with pytest.raises(kma_amd.KmaError):
            h.batch_malloc_det(np.array([64, 4080, 4080], dtype=np.uint32))  # 4080 B: 1 slot at 4 KiB
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 64) as h:
        with pytest.raises(kma_amd.KmaError):
            h.batch_malloc_det(np.array([64], dtype=np.uint32))
