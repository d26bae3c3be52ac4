"""GPU tests of the heap walker / orphan detector (kma_heap_check): heaps
left by real device work check clean, and what they hold is counted
exactly (live blocks of a built graph, page runs of a serialized trace)."""
import numpy as np
import pytest

import kma_amd

pytestmark = pytest.mark.gpu


def _clean(h):
    r = h.check()
    assert r["problems"] == 0 and r["orphans"] == 0, r["first"]
    return r


def test_fresh_and_after_workloads():
    print("heap check: fresh + workloads", flush=True)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 4096, npages=1000) as h:
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
        assert r["live_bytes"] == 32 * st["nodes"] + 16 * st["edges"], (r, st)
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


def test_reference_layout_rejected():
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 12, 64) as h:
        with pytest.raises(kma_amd.KmaError):
            h.check()
