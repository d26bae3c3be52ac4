"""GPU tests of the graph-building application benchmark (the reference's
tb_clTree workload): after one build launch, every node id of the edge list
is in the tree exactly once with exactly its out-edges (degree and sum of
sink ids checked on the device against the host's count of the edge list),
every node allocation is accounted for, and a teardown that frees every node
and link returns the heap to its empty image."""
import numpy as np
import pytest

import kma_amd
from tests import natimg

pytestmark = pytest.mark.gpu


def _check(st, nedges):
    assert st["violations"] == 0 and st["bad"] == 0 and st["nulls"] == 0, st
    assert st["edges"] == nedges and st["links_found"] == nedges and st["link_mallocs"] == nedges, st


@pytest.mark.parametrize("mode", [kma_amd.GRAPH_PLAIN, kma_amd.GRAPH_CACHE, kma_amd.GRAPH_COLLECTIVE])
@pytest.mark.parametrize("nnodes,nedges,nwg", [(3000, 20000, 0), (16, 20000, 64), (100000, 50000, 128)])
def test_graph_build(mode, nnodes, nedges, nwg):
    print("graph mode=%d nodes=%d edges=%d" % (mode, nnodes, nedges), flush=True)
    edges = kma_amd.graph_generate(nnodes, nedges, 11)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 4096) as h:  # 64 MiB
        st = kma_amd.graph_build(edges, mode, heap=h, nwg=nwg, teardown=True)
        _check(st, nedges)
        assert st["nodes"] == len(np.unique(edges))
        # every id kept exactly one node; the rest were freed
        assert st["node_mallocs"] - st["node_frees"] == st["nodes"], st
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


def test_graph_build_on_bump_allocator():
    """The cache form on the bump allocator (the reference's second heap)."""
    print("graph on pma", flush=True)
    edges = kma_amd.graph_generate(3000, 20000, 5)
    with kma_amd.Pma(64 << 20, 16) as p:
        st = kma_amd.graph_build(edges, kma_amd.GRAPH_CACHE, pma=p)
        _check(st, len(edges))
        assert st["node_mallocs"] >= st["nodes"] == len(np.unique(edges))


def test_graph_build_heap_exhaustion_stops_cleanly():
    """A heap too small for the graph: lanes that get NULL stop, nothing
    faults or hangs, and the graph built so far is consistent with what was
    linked (links reached == edges linked)."""
    print("graph exhaustion", flush=True)
    edges = kma_amd.graph_generate(50000, 50000, 3)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 16, nctx=8) as h:  # 256 KiB
        for mode in (kma_amd.GRAPH_PLAIN, kma_amd.GRAPH_COLLECTIVE):
            st = kma_amd.graph_build(edges, mode, heap=h, nwg=32, teardown=True)
            assert st["nulls"] > 0 and st["bad"] == 0, st
            assert st["links_found"] == st["edges"] < len(edges), st
            h.trim()
            assert natimg.NatHeap(h).canonical_problems() == []
