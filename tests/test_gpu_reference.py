"""GPU parity tests: the reference-layout HIP path (kma_ref_device.h) against
the oracle.  Every call goes through libkma.so's C ABI."""
import numpy as np
import pytest

import kma_amd
from oracle import refmodel as rm
from tests import refimg

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("log2,S", [(12, 2), (12, 3), (12, 127), (12, 4096), (14, 4096), (14, 65536)])
def test_init_image_equals_oracle(log2, S):
    _say("ref init log2=%d S=%d" % (log2, S))
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, log2, S) as h:
        img = h.snapshot()
        o = rm.RefHeap(log2, S, base=h.device_ptr)
        assert np.array_equal(img, o.image())


def test_kma_create_defaults_and_rejects():
    with kma_amd.Heap.create(127) as h:  # tb_kma.c:234
        assert h.nbytes == 127 * 4096 + 120
        assert refimg.u64(h.snapshot(0, 32), 0) == h.nbytes
    with pytest.raises(kma_amd.KmaError):
        kma_amd.Heap.create(0)  # kma.c:89-90
    with pytest.raises(kma_amd.KmaError):
        kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 14, 65537)


def test_serial_known_answers():
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 12, 4096) as h:
        out = h.trace_serial([64, 64, 8, 0, 4081, -1, 64])
        assert list(out) == [4100, 4036, 8236, 0, 0, 0, 4100 - 0]  # op 6 reuses slot 62


def _parity(log2, S, ops, pat=None):
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, log2, S) as h:
        out = h.trace_serial(ops, pat)
        img = h.snapshot()
        o = rm.RefHeap(log2, S, base=h.device_ptr)
        oout, _ = o.replay(ops, pat)
        oimg = o.image()
    oout = np.where(oout != 0, oout - o.base, 0).astype(np.uint64)
    assert np.array_equal(out, oout), np.nonzero(out != oout)[0][:10]
    diff = np.nonzero(img != oimg)[0]
    assert diff.size == 0, "image differs at offsets %s" % diff[:16]


@pytest.mark.parametrize("seed", range(4))
def test_serial_trace_parity_4k(seed):
    rng = np.random.default_rng(100 + seed)
    # every class with >= 2 slots (1-slot classes hang the reference)
    sizes = [0, 1, 3, 4, 5, 8, 9, 16, 17, 32, 33, 64, 65, 100, 124, 125, 252, 400, 508,
             1000, 1020, 2040, 4081, 5000]
    ops, _ = refimg.random_trace(rng, 4000, sizes, max_live=300)
    pat = rng.integers(0, 2**32, size=len(ops), dtype=np.uint64).astype(np.uint32)
    _say("ref serial parity 4k seed=%d" % seed)
    _parity(12, 48, ops, pat)


@pytest.mark.parametrize("seed", range(2))
def test_serial_trace_parity_16k_oversubscribed(seed):
    # few superblocks: many NULLs, then frees bring superblocks back
    rng = np.random.default_rng(200 + seed)
    sizes = [8, 64, 128, 129, 508, 1024, 2044, 4092, 8184]
    ops, live = refimg.random_trace(rng, 3000, sizes, free_frac=0.3, max_live=500)
    ops = refimg.free_all(ops, live)
    _say("ref serial parity 16k seed=%d" % seed)
    _parity(14, 12, ops)


def test_serial_plain_small_workload_parity():
    """The plain small workload replayed one work-item at a time (work-item
    0's rounds first): addresses and final image equal the oracle's."""
    nwi, rounds = 1024, 8
    ops, pat = [], []
    for gid in range(nwi):
        for r in range(rounds):
            ops.append(64)
            pat.append(gid)
            ops.append(-(len(ops) - 1) - 1)
            pat.append(0)
    _say("ref serial plain-small parity")
    _parity(12, 4096, np.array(ops, dtype=np.int64), np.array(pat, dtype=np.uint32))


def test_plain_small_concurrent():
    """Plain small workload (16 x 64 work-items, 100 rounds of malloc(64),
    write {gid, round}, read back, free) run concurrently: no NULL, no
    read-back mismatch, and the heap returns to canonical free form."""
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 12, 4096) as h:
        _say("ref plain small concurrent")
        st = h.run(kma_amd.WL_FIXED, 1024, 100, block=64, fixed=64)
        assert st["ops"] == 204800
        assert st["nulls"] == 0 and st["bad"] == 0
        assert refimg.ref_canonical_free(h.snapshot(), 12, 4096) == []


@pytest.mark.parametrize("mode", [kma_amd.WL_REF_TEST, kma_amd.WL_LOWVAR, kma_amd.WL_HIGHVAR])
def test_reference_bench_kernels(mode):
    """kma_test_malloc{,_lowvar,_highvar} (kma.cl:424-526) on the 127-superblock
    heap tb_kma.c:234 creates, 100 iterations (tb_kma.c:376)."""
    with kma_amd.Heap.create(127) as h:
        _say("ref bench kernel mode=%d" % mode)
        st = h.run(mode, 4096, 100, block=64)
        assert st["nulls"] == 0 and st["bad"] == 0
        assert refimg.ref_canonical_free(h.snapshot(), 12, 127) == []


def test_reference_kernel_deadline_bounds_the_launch():
    """kma_test_malloc (kma.cl:424-461) at 131,072 work-items x 100 iterations on
    the 127-superblock heap: the reference protocol needs seconds here (every
    malloc and free CASes one superblock's state word).  With a 300 ms
    deadline the launch ends soon after it, the mallocs that gave up are
    counted and returned NULL, nothing was lost or stranded (claim and
    enqueue give-ups 0), and the heap returns to its canonical free form."""
    with kma_amd.Heap.create(127) as h:
        _say("ref bench kernel with deadline")
        st = h.run(kma_amd.WL_REF_TEST, 131072, 100, block=256, deadline_ms=300)
        assert st["deadline_hits"] > 0, st
        assert st["nulls"] >= st["deadline_hits"] + st["waits"][0], st
        assert st["waits"][1] == 0 and st["waits"][3] == 0, st
        assert st["bad"] == 0
        assert st["gpu_ms"] < 4000, st["gpu_ms"]
        assert refimg.ref_canonical_free(h.snapshot(), 12, 127) == []
        # disarmed afterwards: a small run completes without a give-up
        st = h.run(kma_amd.WL_REF_TEST, 4096, 10, block=64)
        assert st["nulls"] == 0 and st["deadline_hits"] == 0, st


def test_full_occupancy_workload_reference_layout():
    """Mixed classes on the reference layout, concurrently (1 GiB heap at
    16 KiB superblocks, 16,384 work-items, 2 rounds)."""
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 14, 65536) as h:
        _say("ref loguniform concurrent")
        st = h.run(kma_amd.WL_LOGUNI, 16384, 2, block=256, seed=1)
        assert st["bad"] == 0 and st["nulls"] == 0
        assert refimg.ref_canonical_free(h.snapshot(), 14, 65536) == []
