"""GPU tests of the heap walker / orphan detector (kma_heap_check): heaps
left by real device work check clean, and what they hold is counted
exactly (live blocks of a built graph, page runs of a serialized trace)."""
import numpy as np
import pytest

import kma_amd
from tests import natimg

pytestmark = pytest.mark.gpu


def _clean(h):
    r = h.check()
    assert r["problems"] == 0 and r["orphans"] == 0, r["first"]
    return r


@pytest.mark.parametrize("no_partial", [0, 1])
def test_fresh_and_after_workloads(no_partial):
    print("heap check: fresh + workloads, no_partial=%d" % no_partial, flush=True)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 4096, npages=1000, no_partial=no_partial) as h:
        r = _clean(h)
        assert r["sb_free"] == r["sb_total"] == 4096 and r["pages_total"] == 1000
        for form in (kma_amd.FORM_PLAIN, kma_amd.FORM_FUSED, kma_amd.FORM_BATCHED):
            st = h.run(kma_amd.WL_LOGUNI, 65536, 4, form=form, block=256, seed=3)
            # 16 wavefronts share each context here: no reservation may wait
            # out the 0.25 s watchdog on a busy head
            assert st["nulls"] == 0 and st["bad"] == 0 and st["waits"] == [0, 0, 0, 0], st
            assert st["gpu_ms"] < 100, st
            r = _clean(h)
            assert r["live_blocks"] == 0 and r["page_runs"] == 0
            assert r["sb_free"] + r["sb_heads"] == 4096
        h.trim()
        r = _clean(h)
        assert r["sb_free"] == 4096 and r["sb_in_spares"] == 0


@pytest.mark.parametrize("no_partial", [0, 1])
def test_metadata_restored_byte_for_byte(no_partial):
    """All three forms of the full-occupancy workload, then kma_trim: the
    metadata image equals the fresh heap's in every bit outside the
    documented exclusion mask (natimg.restore_mask)."""
    print("heap check: byte comparison, no_partial=%d" % no_partial, flush=True)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 4096, npages=1000, no_partial=no_partial) as h:
        nh = natimg.NatHeap(h)
        init = np.asarray(h.snapshot(0, nh.off_data), dtype=np.uint8).copy()
        for form in (kma_amd.FORM_PLAIN, kma_amd.FORM_FUSED, kma_amd.FORM_BATCHED):
            st = h.run(kma_amd.WL_LOGUNI, 65536, 4, form=form, block=256, seed=7)
            assert st["nulls"] == 0 and st["bad"] == 0, st
        h.trim()
        final = np.asarray(h.snapshot(0, nh.off_data), dtype=np.uint8)
        assert natimg.restored_diff(init, final) == []
        assert nh.canonical_problems() == []


@pytest.mark.parametrize("mode", [kma_amd.GRAPH_PLAIN, kma_amd.GRAPH_CACHE, kma_amd.GRAPH_COLLECTIVE])
def test_graph_live_blocks_counted(mode):
    """A graph built without teardown holds one block per node and one per
    edge; after the teardown frees them, none."""
    print("heap check: graph mode=%d" % mode, flush=True)
    edges = kma_amd.graph_generate(3000, 20000, 7)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 4096) as h:
        st = kma_amd.graph_build(edges, mode, heap=h)
        assert st["violations"] == 0 and st["nulls"] == 0, st
        r = _clean(h)
        assert r["live_blocks"] == st["nodes"] + st["edges"], (r, st)
        # 48-B nodes live in the 64-B class of the default table
        assert r["live_bytes"] == 64 * st["nodes"] + 16 * st["edges"], (r, st)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 4096) as h:
        st = kma_amd.graph_build(edges, mode, heap=h, teardown=True)
        r = _clean(h)
        assert r["live_blocks"] == 0, (r, st)


def test_page_runs_counted():
    """Runs of 1, 3, 64 and 5 pages and two small blocks left live: the walk
    finds each run with its length, and both blocks."""
    print("heap check: page runs", flush=True)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 64, nctx=8, npages=512) as h:
        out = h.trace_serial([1 << 14, 3 << 14, 64 << 14, 5 << 14, 24, 24, 2 << 14, -7])
        assert np.count_nonzero(out[:7]) == 7
        r = _clean(h)
        assert r["page_runs"] == 4 and r["pages_used"] == 1 + 3 + 64 + 5, r
        assert r["live_blocks"] == 2, r
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 64, nctx=8, npages=512) as h:
        out = h.trace_serial([1 << 14, 3 << 14, 64 << 14, 5 << 14, -2, -4])
        r = _clean(h)
        assert r["page_runs"] == 2 and r["pages_used"] == 65, r


@pytest.mark.parametrize("mode", [kma_amd.WL_REF_TEST, kma_amd.WL_LOWVAR, kma_amd.WL_HIGHVAR])
def test_reference_layout_walk_after_bench_kernels(mode):
    """The reference's bench kernels (kma.cl:424-526) on kma_create(127) at
    32,768 work-items x 20 iterations in one launch (as tb_kma.c runs them),
    then the reference-layout walk: the free queue ends at its tail, no
    superblock is lost, orphaned or miscounted, and no give-up was counted.
    Every block is free again except, in _lowvar and _highvar, work-item 0's
    small block when its malloc(4000) (a 1-slot class, which needs a whole
    superblock) got NULL: the kernels return without freeing it
    (kma.cl:484-486, 516-518)."""
    print("heap check: reference layout mode=%d" % mode, flush=True)
    with kma_amd.Heap.create(127) as h:
        r = h.check()
        assert r["problems"] == 0 and r["sb_free"] == 126 and r["sb_unqueued"] == 1, r
        st = h.run(mode, 32768, 20, block=256, form=kma_amd.FORM_FUSED)
        assert st["bad"] == 0 and sum(st["waits"]) == 0 and st["deadline_hits"] == 0, st
        r = h.check()
        assert r["problems"] == 0, r["first"]
        leaked = 0 if mode == kma_amd.WL_REF_TEST else 1
        assert r["live_blocks"] <= leaked and r["orphans"] == 0, r
        if r["live_blocks"]:
            assert st["nulls"] > 0 and r["live_bytes"] <= 64, (r, st)
        else:
            assert r["sb_heads"] == 0 and r["sb_free"] == 126, r
        assert r["sb_total"] == 127 and r["counters"][0] == 0, r  # deadline disarmed


def test_reference_layout_walk_with_live_blocks():
    """A serialized trace leaving blocks live in four classes: the walk
    counts exactly the blocks the trace left and the superblocks it used."""
    print("heap check: reference layout live blocks", flush=True)
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 12, 64) as h:
        ops = [64] * 70 + [8] * 10 + [1020] * 9 + [2040] * 3 + [-1, -2, -71]
        out = h.trace_serial(ops)
        assert np.count_nonzero(out[:92]) == 92
        r = h.check()
        assert r["problems"] == 0, r["first"]
        assert r["live_blocks"] == 92 - 3, r
        assert r["live_bytes"] == 68 * 64 + 9 * 8 + 9 * 1020 + 3 * 2040, r
