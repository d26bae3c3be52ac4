"""Host side of the graph-building benchmark: the edge-list generator and
the reference's data-set text format (tb_clTree.c:350-389).  No GPU."""
import numpy as np
import pytest

import kma_amd


def test_generator_is_deterministic_and_uses_all_ids():
    a = kma_amd.graph_generate(1000, 20000, 7)
    b = kma_amd.graph_generate(1000, 20000, 7)
    c = kma_amd.graph_generate(1000, 20000, 8)
    assert a.shape == (20000, 2) and a.dtype == np.uint32
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # 40,000 draws over 1,000 ids: every id appears, and no other value
    assert len(np.unique(a)) == 1000
    assert len(np.unique(kma_amd.graph_generate(1000, 20000, 9))) == 1000
    assert set(np.unique(a).tolist()) == set(np.unique(c).tolist())


def test_dataset_text_round_trip(tmp_path):
    e = kma_amd.graph_generate(50, 300, 3)
    p = tmp_path / "edges.txt"
    kma_amd.graph_write(p, e)
    lines = p.read_text().splitlines()
    assert lines[0] == "300" and lines[1] == "{%d,%d}" % (int(e[0, 0]), int(e[0, 1]))
    assert np.array_equal(kma_amd.graph_read(p), e)


def test_dataset_malformed_line_rejected(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("2\n{1,2}\n3,4\n")
    with pytest.raises(ValueError):
        kma_amd.graph_read(p)


def test_build_argument_checks():
    e = kma_amd.graph_generate(10, 10, 1)
    with pytest.raises(kma_amd.KmaError):
        kma_amd.graph_build(e, kma_amd.GRAPH_PLAIN)  # neither heap nor bump allocator
    with pytest.raises(kma_amd.KmaError):
        kma_amd.graph_build(e, 7)
