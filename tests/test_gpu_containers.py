"""Bump allocator, pointer queue, array list and binary search tree on the GPU
(include/kma_containers.h), checked against closed forms and the reference
testbenches' validators (tb_clQueue.c:126-157: the walk from head ends at
the tail, counts every item once, has no cycle)."""
import numpy as np
import pytest

import kma_amd
from tests.natimg import NatHeap

pytestmark = pytest.mark.gpu
NULL = np.uint64(0xFFFFFFFFFFFFFFFF)


# ---------------- bump allocator (pma.cl:32-48) ----------------
@pytest.mark.parametrize("align", [1, 16])
def test_pma_serial_closed_form(align):
    """k-th serialized call returns base + sum of the earlier sizes (rounded
    to the alignment); once a request does not fit, it and every later one
    fail, and the head still counts them (pma.cl:37-39)."""
    print("pma serial align=%d" % align, flush=True)
    rng = np.random.default_rng(1)
    sizes = rng.integers(1, 4097, size=1500).astype(np.uint32)
    rs = (sizes.astype(np.int64) + align - 1) // align * align
    pre = np.concatenate([[0], np.cumsum(rs)[:-1]])
    cap = int(pre[1000])  # the first 1000 fit exactly
    with kma_amd.Pma(cap, align) as p:
        out = p.run(sizes, serial=True)
        ok = pre + rs <= cap
        assert np.array_equal(out[ok], pre[ok].astype(np.uint64))
        assert (out[~ok] == NULL).all()
        assert ok.sum() == 1000 and ok[:1000].all()
        assert p.head == int(rs.sum())
        p.reset()
        assert p.head == 0
        assert np.array_equal(p.run(sizes[:10], serial=True), pre[:10].astype(np.uint64))


def test_pma_parallel_tiles_region():
    """Concurrent requests (one atomic per wavefront): the blocks tile
    [0, total) with no gap or overlap, lanes of a wavefront in lane order."""
    print("pma parallel", flush=True)
    rng = np.random.default_rng(2)
    n = 200000
    sizes = rng.integers(1, 300, size=n).astype(np.uint32)
    total = int(sizes.astype(np.int64).sum())
    with kma_amd.Pma(total + 1000, 1) as p:
        out = p.run(sizes)
        assert (out != NULL).all()
        order = np.argsort(out)
        o, s = out[order].astype(np.int64), sizes[order].astype(np.int64)
        assert o[0] == 0 and (o[1:] == o[:-1] + s[:-1]).all() and o[-1] + s[-1] == total
        w = out.reshape(-1, 64).astype(np.int64)
        sw = sizes.reshape(-1, 64).astype(np.int64)
        assert (w[:, 1:] == w[:, :-1] + sw[:, :-1]).all()
        assert p.head == total


def test_pma_parallel_exhaustion():
    print("pma exhaustion", flush=True)
    rng = np.random.default_rng(3)
    sizes = rng.integers(16, 1024, size=65536).astype(np.uint32)
    cap = int(sizes.astype(np.int64).sum() // 3)
    with kma_amd.Pma(cap, 16) as p:
        out = p.run(sizes)
        ok = out != NULL
        rs = (sizes.astype(np.int64) + 15) // 16 * 16
        ends = out[ok].astype(np.int64) + rs[ok]
        assert ends.max() <= cap
        o = np.sort(out[ok].astype(np.int64))
        r = rs[ok][np.argsort(out[ok])]
        assert (o[1:] >= o[:-1] + r[:-1]).all()  # no overlap
        assert p.head == int(rs.sum())  # failed requests still advance the head


def test_pma_rejects():
    with pytest.raises(kma_amd.KmaError):
        kma_amd.Pma(0)
    with pytest.raises(kma_amd.KmaError):
        kma_amd.Pma(1024, 3)


# ---------------- pointer queue (clQueue.cl:40-162) ----------------
def test_queue_enqueue_test_kernel():
    """clQueue_test_enqueue: every work-item enqueues its item; the queue
    then holds each item exactly once."""
    print("queue enqueue", flush=True)
    n = 100000
    with kma_amd.Queue(1 << 17) as q:
        base, _ = q.run(0, n)
        items = q.items()
        assert len(items) == n
        idx = (items - np.uint64(base)) // np.uint64(16)
        assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.uint64))


def test_queue_shuffle_test_kernel():
    """clQueue_test_dequeue: enqueue, then 10 x (dequeue, enqueue again):
    nothing lost, nothing duplicated."""
    print("queue shuffle", flush=True)
    n = 65536
    with kma_amd.Queue(2 * n) as q:
        base, _ = q.run(1, n, rounds=10)
        items = q.items()
        assert len(items) == n
        idx = (items - np.uint64(base)) // np.uint64(16)
        assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.uint64))


def test_queue_fifo_serial():
    print("queue fifo", flush=True)
    n = 5000
    with kma_amd.Queue(8192) as q:
        _, order = q.run(2, n)
        assert np.array_equal(order, np.arange(n))
        assert len(q.items()) == 0
        # the ring wraps: a second lap over the same slots
        _, order = q.run(2, n)
        assert np.array_equal(order, np.arange(n))


def test_queue_capacity_checked():
    with kma_amd.Queue(64) as q:
        q.run(0, 64)
        assert len(q.items()) == 64
        with pytest.raises(kma_amd.KmaError):
            q.run(0, 65)
        q.reset()
        assert len(q.items()) == 0


# ------- unbounded linked queue (clQueue.cl:40-162, tb_clQueue.c:126-157) -------
def _valid(rep, n):
    """The reference's validator: the chain from head ends, its last node is
    the one tail names, and it holds each item once (plus the stub)."""
    assert rep["lost"] == 0 and rep["outside"] == 0 and rep["repeats"] == 0, rep
    assert rep["cycle"] == 0 and rep["tail_ok"] == 1, rep
    assert rep["items"] == n and rep["stubs"] == 1, rep


def test_lqueue_enqueue_test_kernel():
    """clQueue_test_enqueue at 100,000 work-items: every item linked once,
    no capacity of any kind to size."""
    print("linked queue enqueue", flush=True)
    with kma_amd.LinkedQueue() as q:
        rep, _ = q.run(0, 100000)
        _valid(rep, 100000)


@pytest.mark.parametrize("n", [512, 16384])
def test_lqueue_shuffle_test_kernel(n):
    """clQueue_test_dequeue (clQueue.cl:140-162): every work-item enqueues its
    item, then 10 x (dequeue, enqueue again): nothing lost, nothing
    duplicated (512 work-items: the reference's GLOBAL_SIZE,
    tb_clQueue.c:31-34)."""
    print("linked queue shuffle %d" % n, flush=True)
    with kma_amd.LinkedQueue() as q:
        rep, _ = q.run(1, n, rounds=10)
        _valid(rep, n)


def test_lqueue_fifo_serial():
    """One work-item: items come out in the order they went in, and the
    queue is empty (only its stub left) afterwards; twice, so the stub goes
    round again."""
    print("linked queue fifo", flush=True)
    with kma_amd.LinkedQueue() as q:
        for _ in range(2):
            rep, order = q.run(2, 5000)
            assert np.array_equal(order, np.arange(5000))
            _valid(rep, 0)
        rep, order = q.run(2, 0)
        _valid(rep, 0)


# ---------------- array list (clArrayList.cl:32-216) ----------------
def _pages_and_counts(h, lst):
    pages = lst.pages()
    snap = h.snapshot()
    hdr = [np.frombuffer(snap[int(p):int(p) + 24].tobytes(), dtype="<u8") for p in pages]
    return pages, hdr, snap


def test_list_grow_local_closed_form():
    """Each work-group takes one page for the sum of its requests; thread i
    gets data + (exclusive prefix of objs within its work-group) * obj_size
    (clArrayList.cl:156-170), NULL if it asked for none."""
    print("list grow_local", flush=True)
    rng = np.random.default_rng(4)
    nwi, block, osz = 4096, 256, 16
    objs = rng.integers(0, 6, size=nwi).astype(np.uint32)
    # a work-group's page (~10 KiB) is above the largest class: page region
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 2048, npages=512) as h, kma_amd.ArrayList(h, osz) as lst:
        out = lst.grow(objs, block=block, local=True)
        pages, hdr, snap = _pages_and_counts(h, lst)
        assert len(pages) == nwi // block
        by_page = {}
        for g in range(nwi // block):
            o = objs[g * block:(g + 1) * block].astype(np.int64)
            pre = np.concatenate([[0], np.cumsum(o)[:-1]])
            got = out[g * block:(g + 1) * block].astype(np.int64)
            first = int(np.nonzero(o)[0][0])
            data = got[first] - pre[first] * osz
            assert ((got == data + pre * osz) | (o == 0)).all()
            assert (got[o == 0] == 0).all()
            by_page[data - 24] = int(o.sum())
        for p, hd in zip(pages, hdr):
            assert by_page[int(p)] == int(hd[1]) == int(hd[2]) and int(hd[0]) == 0
        # every object carries its owner's pattern (thread i, object k)
        words = np.frombuffer(snap.tobytes(), dtype="<u4")
        for i in np.nonzero(objs)[0][:500]:
            for k in range(int(objs[i])):
                assert words[(int(out[i]) + k * osz) // 4] == (int(i) ^ k)


def test_list_get_and_clear():
    print("list get/clear", flush=True)
    rng = np.random.default_rng(5)
    nwi, osz = 2048, 8
    objs = rng.integers(1, 4, size=nwi).astype(np.uint32)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 2048) as h, kma_amd.ArrayList(h, osz) as lst:
        lst.grow(objs, block=128, local=True)
        lst.grow(objs[:300], block=64, local=False)  # one page per work-item
        pages, hdr, _ = _pages_and_counts(h, lst)
        counts = [int(x[1]) for x in hdr]
        assert len(pages) == nwi // 128 + 300
        total = sum(counts)
        assert total == int(objs.sum()) + int(objs[:300].sum())
        expect = []
        for p, c in zip(pages, counts):
            expect += [int(p) + 24 + k * osz for k in range(c)]
        idx = np.arange(total + 3, dtype=np.uint64)
        got = lst.get(idx)
        assert np.array_equal(got[:total], np.array(expect, dtype=np.uint64))
        assert (got[total:] == 0).all()
        lst.clear(nwi=512)
        assert len(lst.pages()) == 0
        h.trim()
        assert NatHeap(h).canonical_problems() == []


# ---------------- binary search tree (clTree.cl:27-74) ----------------
def test_tree_concurrent_insert_and_lookup():
    """Concurrent adds with duplicate keys: exactly one node per distinct key
    is linked, the in-order walk is the sorted distinct keys, lookups find
    exactly the inserted keys; the heap is empty again afterwards."""
    print("tree", flush=True)
    rng = np.random.default_rng(6)
    keys = rng.integers(0, 30000, size=40000).astype(np.uint32)
    distinct = np.unique(keys)
    absent = np.setdiff1d(np.arange(30000, dtype=np.uint32), distinct)[:2000]
    query = np.concatenate([distinct, absent, np.array([40000, 0xFFFFFFFF], dtype=np.uint32)])
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 1024) as h:
        added, found, inorder, bad = kma_amd.tree_run(h, keys, query)
        assert not (added == 2).any(), "heap could not serve a node"
        assert int(added.sum()) == len(distinct)
        for k in distinct[:200]:
            assert added[keys == k].sum() == 1
        assert bad == 0
        assert np.array_equal(inorder, distinct)
        assert found[:len(distinct)].all() and not found[len(distinct):].any()
        h.trim()
        assert NatHeap(h).canonical_problems() == []
