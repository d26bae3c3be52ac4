"""Batched request streams on the GPU (kma_stream_run): the per-launch form
(one launch per batch, per-lane API) and the persistent form (one
graph-captured launch, work-group collective API).

Checks, for every form: the request accounting identities, the pattern each
malloc writes into its block (read back by the free), per-block invariants
(in-heap, aligned, class-correct), that no two blocks sharing memory were
live at the same time (from per-request device wall-clock ticks), and the
canonical empty-heap image after the draining run."""
import numpy as np
import pytest

import kma_amd
from tests import natimg
from tests.natimg import NatHeap

pytestmark = pytest.mark.gpu

CLASSES = [16 << i for i in range(11)]  # 16 B .. 16 KiB


def stream_heap(sblocks=2048, npages=4096, sb_log2=14, no_partial=0):
    return kma_amd.Heap(layout=kma_amd.LAYOUT_NATIVE, sb_log2=sb_log2, sblocks=sblocks,
                        classes=[c for c in CLASSES if c <= (1 << sb_log2)], page_log2=14,
                        npages=npages, no_partial=no_partial)


def extents(nh, offs, sizes):
    """Footprint of each block: its class for superblock blocks, its page
    run for page blocks."""
    psz = 1 << nh.page_log2
    ln = np.zeros(len(offs), dtype=np.int64)
    for i, (o, s) in enumerate(zip(offs, sizes)):
        if s <= nh.max_small:
            ln[i] = next(c for c in nh.classes if c >= s)
        else:
            ln[i] = (int(s) + psz - 1) // psz * psz
    return ln


def block_problems(nh, offs, sizes):
    probs = []
    data_end = nh.off_data + (nh.nsb << nh.sb_log2)
    psz = 1 << nh.page_log2
    for off, size in zip(offs, sizes):
        off, size = int(off), int(size)
        if off == 0:  # NULL
            continue
        if size <= nh.max_small:
            cls = next(c for c in nh.classes if c >= size)
            inner = (off - nh.off_data) & ((1 << nh.sb_log2) - 1)
            if not nh.off_data <= off < data_end or inner % cls or inner + cls > (1 << nh.sb_log2):
                probs.append("small block %#x (size %d) misplaced" % (off, size))
        else:
            if not nh.off_pages <= off < nh.off_pages + (nh.npages << nh.page_log2) or (off - nh.off_pages) % psz:
                probs.append("page block %#x misplaced" % off)
            k = (size + psz - 1) // psz
            if (off - nh.off_pages) // psz % 64 + k > 64:
                probs.append("page run %#x crosses a bitmap word" % off)
    return probs


def lifetime_overlaps(req, res, ticks, nh):
    """Blocks that share memory must not have been live at the same time:
    block j is live at least over [tick of its malloc's return, tick of the
    free that names it]; two such intervals may touch (10 ns ticks) but not
    overlap."""
    n = len(req)
    isf = (req >> 31) == 1
    tgt = (req & 0x7FFFFFFF).astype(np.int64)
    free_tick = np.full(n, -1, dtype=np.int64)
    free_tick[tgt[isf]] = ticks[isf].astype(np.int64)
    m = np.nonzero((~isf) & (res != 0))[0]
    if len(m) == 0:
        return []
    offs = res[m].astype(np.int64)
    ln = extents(nh, offs, req[m])
    t0 = ticks[m].astype(np.int64)
    t1 = free_tick[m]
    assert (t1 >= 0).all(), "a block was never freed"
    order = np.argsort(offs, kind="stable")
    probs = []
    # sweep by address: compare each block with later-starting blocks that
    # begin inside it
    active = []
    for i in order:
        a, e = offs[i], offs[i] + ln[i]
        active = [k for k in active if offs[k] + ln[k] > a]
        for k in active:
            if max(t0[i], t0[k]) < min(t1[i], t1[k]):
                probs.append("requests %d and %d share memory while both live" % (m[i], m[k]))
                if len(probs) > 5:
                    return probs
        active.append(i)
    return probs


def check_run(st, info):
    assert st["ops"] == info["nreq"], st
    assert st["mallocs"] + st["nulls"] == info["nmalloc"], st
    assert st["frees"] + st["free_nulls"] == info["nfree"], st
    assert st["free_nulls"] == st["nulls"], st  # the drain frees every NULL exactly once
    assert st["bad"] == 0 and st["wait_timeouts"] == 0, st
    assert st["double_frees"] == 0 and st["bad_frees"] == 0, st
    assert st["waits"] == [0, 0, 0, 0], st


@pytest.mark.parametrize("form,nwg,per_thread", [
    (kma_amd.FORM_PLAIN, 0, 0),
    (kma_amd.FORM_BATCHED, 64, 1),
    (kma_amd.FORM_BATCHED, 96, 4),
])
def test_stream_invariants(form, nwg, per_thread):
    print("stream invariants form=%d nwg=%d per_thread=%d" % (form, nwg, per_thread), flush=True)
    with stream_heap() as h, kma_amd.Stream(48, 4096, seed=3) as s:
        st = s.run(h, form=form, nwg=nwg, per_thread=per_thread, variant=kma_amd.VAR_LOG)
        check_run(st, s.info)
        assert st["nulls"] == 0, st  # the heap holds the whole live set
        req, _ = s.trace()
        res, ticks = s.results(), s.times()
        nh = NatHeap(h)
        mal = (req >> 31) == 0
        assert not block_problems(nh, res[mal], req[mal])
        assert not lifetime_overlaps(req, res, ticks, nh)
        h.trim()
        assert nh.canonical_problems() == []


@pytest.mark.parametrize("no_partial", [0, 1])
@pytest.mark.parametrize("form", [kma_amd.FORM_PLAIN, kma_amd.FORM_BATCHED])
def test_stream_oversubscribed(form, no_partial):
    """A heap far smaller than the live set: NULLs, and still every
    invariant and the empty-heap image afterwards; with and without reuse
    of partly free superblocks."""
    print("stream oversubscribed form=%d no_partial=%d" % (form, no_partial), flush=True)
    with stream_heap(sblocks=48, npages=96, no_partial=no_partial) as h, kma_amd.Stream(40, 4096, seed=5) as s:
        st = s.run(h, form=form, nwg=128, per_thread=2, variant=kma_amd.VAR_LOG)
        check_run(st, s.info)
        assert st["nulls"] > 0 and st["large_nulls"] > 0, st
        req, _ = s.trace()
        res, ticks = s.results(), s.times()
        nh = NatHeap(h)
        mal = (req >> 31) == 0
        assert not block_problems(nh, res[mal], req[mal])
        assert not lifetime_overlaps(req, res, ticks, nh)
        h.trim()
        assert nh.canonical_problems() == []


def test_stream_large_page_region():
    """Page runs from a 20,000-page region (several summary words per level,
    the top-level summary in use) under the persistent form: invariants and
    the empty-heap image afterwards."""
    print("stream large page region", flush=True)
    with stream_heap(sblocks=512, npages=20000) as h, kma_amd.Stream(48, 4096, seed=13) as s:
        st = s.run(h, form=kma_amd.FORM_BATCHED, nwg=128, per_thread=2, variant=kma_amd.VAR_LOG)
        check_run(st, s.info)
        assert st["large"] > 0, st
        req, _ = s.trace()
        res, ticks = s.results(), s.times()
        nh = NatHeap(h)
        mal = (req >> 31) == 0
        assert not block_problems(nh, res[mal], req[mal])
        assert not lifetime_overlaps(req, res, ticks, nh)
        h.trim()
        assert nh.canonical_problems() == []


def test_stream_restores_metadata_byte_for_byte():
    """Replays of a draining stream with page runs and NULLs, then kma_trim:
    the metadata image equals the fresh heap's in every bit outside the
    documented exclusion mask (natimg.restore_mask), page bitmap and
    page-window summaries included."""
    print("stream byte comparison", flush=True)
    with stream_heap(sblocks=96, npages=256) as h, kma_amd.Stream(40, 4096, seed=5) as s:
        nh = NatHeap(h)
        init = np.asarray(h.snapshot(0, nh.off_data), dtype=np.uint8).copy()
        st = s.run(h, form=kma_amd.FORM_BATCHED, nwg=128, per_thread=2, warmup=1, reps=2)
        check_run(st, s.info)
        assert st["large"] > 0 and st["nulls"] > 0, st
        h.trim()
        final = np.asarray(h.snapshot(0, nh.off_data), dtype=np.uint8)
        assert natimg.restored_diff(init, final) == []


def test_stream_replays_return_heap_to_empty():
    """Warm-up plus timed replays of the graph: each replay starts from the
    state the previous draining replay left."""
    print("stream replays", flush=True)
    with stream_heap() as h, kma_amd.Stream(64, 8192, seed=11) as s:
        st = s.run(h, form=kma_amd.FORM_BATCHED, nwg=256, per_thread=2, warmup=1, reps=3)
        check_run(st, s.info)
        assert st["launches"] == 1
        h.trim()
        assert NatHeap(h).canonical_problems() == []
        st2 = s.run(h, form=kma_amd.FORM_PLAIN, warmup=1, reps=2)
        check_run(st2, s.info)
        assert st2["launches"] == s.info["nbatch"]
        h.trim()
        assert NatHeap(h).canonical_problems() == []


def test_stream_autotune_small():
    """The autotuner's pick satisfies its own constraints and runs clean."""
    print("stream autotune", flush=True)
    with kma_amd.Stream(48, 4096, seed=9) as s:
        r = s.autotune(256 << 20, 14, null_limit=0, reps=1)
        assert len(r["trials"]) == 97
        assert all(t["violations"] == 0 for t in r["trials"]), r["trials"]
        best = max(r["trials"], key=lambda t: t["ops_per_s"])
        assert r["ops_per_s"] == best["ops_per_s"]
        assert r["config"]["sb_log2"] == best["sb_log2"] and r["nwg"] == best["nwg"] and r["agg"] == best["agg"]
        assert r["config"]["no_partial"] == best["no_partial"]
        assert {t["agg"] for t in r["trials"]} == {0x00, 0x10, 0x20, 0x30}
        # the last trial repeats the best of the first 96 on a heap without
        # partial reuse
        first = max(r["trials"][:96], key=lambda t: t["ops_per_s"])
        last = r["trials"][96]
        assert last["no_partial"] == 1 and all(t["no_partial"] == 0 for t in r["trials"][:96])
        assert {k: last[k] for k in ("sb_log2", "table", "agg", "nwg", "per_thread")} == \
            {k: first[k] for k in ("sb_log2", "table", "agg", "nwg", "per_thread")}
        with kma_amd.Heap.from_config(r["cfg"]) as h:
            st = s.run(h, form=kma_amd.FORM_BATCHED, nwg=r["nwg"], per_thread=r["per_thread"], variant=r["agg"])
            check_run(st, s.info)
            h.trim()
            assert NatHeap(h).canonical_problems() == []
