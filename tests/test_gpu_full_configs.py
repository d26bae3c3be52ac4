"""The batched-stream and serving configurations at their stated sizes, on
the GPU, checked by properties that do not depend on the size.

Batched persistent stream: 4,096 batches, batch sizes Zipf(1.2) over
1..65,536 requests, sizes log-uniform 16 B..64 KiB (>16 KiB on 16 KiB
pages), 40 % frees of earlier batches' live blocks, seed 3, a drain batch
at the end, on a 4 GiB heap split between superblocks and pages by the
trace's demand.  Both forms (one launch per batch; one graph-captured
persistent launch).  Checks: the accounting identities, the patterns read
back by the frees (counted on the device), every block in its region,
aligned and inside its superblock or page word, no two blocks sharing
memory while both live (from per-request device ticks, vectorised: blocks
at one address, class epochs of each superblock, and every page), and the
empty-heap metadata after the drain.

Serving loop: Poisson arrivals at 2,000 jobs/ms, log-normal lengths (mu 6,
sigma 1 steps, capped at 8,192), one 16 KiB block per live job per step,
64 GiB heap, admission below 90 % use, 10 s, seed 4.  Checks: the
accounting identities, every sampled completed job's result equal to its
closed form, three jobs run alone equal to their result in the mix
(tolerance: exact), and the empty-heap metadata afterwards."""
import numpy as np
import pytest

import kma_amd
from tests.natimg import NatHeap

pytestmark = pytest.mark.gpu

STREAM = dict(nbatch=4096, max_batch=65536, zipf=1.2, free_pct=40, min_size=16, max_size=65536,
              page_min=16384, seed=3)
STREAM_HEAP = 4 << 30


def block_problems(nh, offs, sizes):
    """Blocks outside their region, misaligned for their class, spilling
    past their superblock, or page runs crossing a 64-page bitmap word."""
    offs = offs.astype(np.int64)
    sizes = sizes.astype(np.int64)
    classes = np.array(nh.classes, dtype=np.int64)
    sb = 1 << nh.sb_log2
    small = (offs != 0) & (sizes <= nh.max_small)
    large = (offs != 0) & (sizes > nh.max_small)
    cls = classes[np.searchsorted(classes, sizes[small])]
    o = offs[small]
    inner = (o - nh.off_data) & (sb - 1)
    data_end = nh.off_data + (nh.nsb << nh.sb_log2)
    bad_small = (o < nh.off_data) | (o >= data_end) | (inner % cls != 0) | (inner + cls > sb)
    psz = 1 << nh.page_log2
    po = offs[large] - nh.off_pages
    k = (sizes[large] + psz - 1) // psz
    bad_large = (po < 0) | (po >= nh.npages * psz) | (po % psz != 0) | ((po // psz) % 64 + k > 64)
    return int(np.count_nonzero(bad_small)), int(np.count_nonzero(bad_large))


def _overlaps_by_key(key, t0, t1):
    """Pairs of consecutive lifetimes (by start tick) on one key that
    overlap: a lifetime may end at the tick the next begins, not after."""
    o = np.lexsort((t0, key))
    k, a, b = key[o], t0[o], t1[o]
    return int(np.count_nonzero((k[1:] == k[:-1]) & (a[1:] < b[:-1])))


def lifetime_problems(nh, req, res, ticks):
    """Blocks that shared memory while both live.  A block lives from the
    tick its malloc returned to the tick the free naming it was issued.
    Small blocks: lifetimes at one address must not overlap, and in each
    superblock the runs of one class (ordered by malloc tick) must end
    before the next class's first block starts (a superblock changes class
    only through the pool); page runs: no two lifetimes on one page."""
    n = len(req)
    isf = (req >> 31) == 1
    tgt = (req & 0x7FFFFFFF).astype(np.int64)
    free_tick = np.full(n, -1, dtype=np.int64)
    free_tick[tgt[isf]] = ticks[isf].astype(np.int64)
    m = (~isf) & (res != 0)
    offs = res[m].astype(np.int64)
    sizes = req[m].astype(np.int64)
    t0 = ticks[m].astype(np.int64)
    t1 = free_tick[m]
    assert (t1 >= 0).all(), "a block was never freed"
    assert (t1 >= t0).all(), "a free was issued before its malloc returned"
    small = sizes <= nh.max_small
    out = {"same_address": _overlaps_by_key(offs[small], t0[small], t1[small])}
    classes = np.array(nh.classes, dtype=np.int64)
    sbi = (offs[small] - nh.off_data) >> nh.sb_log2
    cls = np.searchsorted(classes, sizes[small])
    o = np.lexsort((t0[small], sbi))
    s_o, c_o, a_o, b_o = sbi[o], cls[o], t0[small][o], t1[small][o]
    first = np.nonzero(np.r_[True, (s_o[1:] != s_o[:-1]) | (c_o[1:] != c_o[:-1])])[0]
    e_sb, e_t0 = s_o[first], a_o[first]
    e_t1 = np.maximum.reduceat(b_o, first)
    out["class_epochs"] = int(np.count_nonzero((e_sb[1:] == e_sb[:-1]) & (e_t0[1:] < e_t1[:-1])))
    psz = 1 << nh.page_log2
    lg = ~small
    p0 = (offs[lg] - nh.off_pages) // psz
    k = (sizes[lg] + psz - 1) // psz
    starts = np.repeat(np.cumsum(k) - k, k)
    pages = np.repeat(p0, k) + (np.arange(int(k.sum())) - starts)
    out["pages"] = _overlaps_by_key(pages, np.repeat(t0[lg], k), np.repeat(t1[lg], k))
    out["blocks"] = int(m.sum())
    out["page_runs"] = int(lg.sum())
    return out


@pytest.fixture(scope="module")
def stream():
    print("full-size stream: generating", flush=True)
    with kma_amd.Stream(**STREAM) as s:
        assert s.info["nbatch"] == 4097 and s.info["nreq"] > 7_000_000, s.info
        yield s


@pytest.mark.parametrize("form", [kma_amd.FORM_BATCHED, kma_amd.FORM_PLAIN])
def test_batched_stream_full_size(stream, form):
    s = stream
    print("full-size stream form=%d: %d requests in %d batches" % (form, s.info["nreq"], s.info["nbatch"]),
          flush=True)
    cfg = s.heap_config(STREAM_HEAP, 14, kma_amd.TABLE_POW2, 14)
    with kma_amd.Heap.from_config(cfg) as h:
        st = s.run(h, form=form, nwg=1024, per_thread=1, variant=kma_amd.VAR_LOG)
        info = s.info
        assert st["ops"] == info["nreq"], st
        assert st["mallocs"] + st["nulls"] == info["nmalloc"], st
        assert st["frees"] + st["free_nulls"] == info["nfree"], st
        assert st["free_nulls"] == st["nulls"], st
        assert st["bad"] == 0 and st["wait_timeouts"] == 0, st
        assert st["double_frees"] == 0 and st["bad_frees"] == 0 and st["waits"] == [0, 0, 0, 0], st
        # the live set outgrows the heap: NULLs, of small blocks and page runs
        assert st["nulls"] > 0 and st["large_nulls"] > 0, st
        req, _ = s.trace()
        res, ticks = s.results(), s.times()
        nh = NatHeap(h)
        mal = (req >> 31) == 0
        assert block_problems(nh, res[mal], req[mal]) == (0, 0)
        lp = lifetime_problems(nh, req, res, ticks)
        print("full-size stream form=%d: %s" % (form, lp), flush=True)
        assert lp["same_address"] == 0 and lp["class_epochs"] == 0 and lp["pages"] == 0, lp
        assert lp["blocks"] == st["mallocs"] and lp["page_runs"] > 0, (lp, st)
        h.trim()
        assert nh.canonical_problems() == []


SERVE = dict(rate_per_ms=2000.0, len_mu=6.0, len_sigma=1.0, duration_ms=10000.0, max_len=8192,
             block_bytes=16384, watermark_pct=90, max_live=32768, seed=4)


def expected(spec, jobs):
    ln = kma_amd.serve_lengths(spec, jobs)
    out = np.zeros(len(jobs), dtype=np.uint64)
    for i, (j, L) in enumerate(zip(jobs, ln)):
        with np.errstate(over="ignore"):
            out[i] = kma_amd.serve_pattern(int(j), np.arange(int(L), dtype=np.uint64)).sum(dtype=np.uint64)
    return out


def test_serving_loop_full_size():
    spec = kma_amd.serve_spec(**SERVE)
    arr, _ = kma_amd.serve_trace(spec)
    print("full-size serving loop: %d jobs, 64 GiB heap, 10 s" % len(arr), flush=True)
    with kma_amd.Heap(layout=kma_amd.LAYOUT_NATIVE, sb_log2=20, sblocks=65536, classes=[16384]) as h:
        st, res = kma_amd.serve_run(h, spec, nresults=len(arr))
        print("full-size serving loop: %s" % {k: v for k, v in st.items() if not isinstance(v, float)},
              flush=True)
        assert not st["aborted"] and st["bad"] == 0, st
        assert st["mallocs"] == st["frees"] + st["drain_frees"], st
        assert st["admitted"] == st["done"] + st["drained"] + st["preempted"], st
        assert st["jobs"] == len(arr) and 9500 < st["run_ms"] < 10500, st
        # projected admission: (almost) no preemption, and the projection of
        # every step without stalls equals the blocks held after it
        assert st["admit"] == kma_amd.ADMIT_PROJECTED and st["proj_mismatch"] == 0, st
        assert st["preempted"] <= 0.1 * st["admitted"], st
        assert st["proj_checked"] >= 0.9 * st["steps"], st
        done = np.nonzero(res)[0].astype(np.uint64)
        assert len(done) == st["done"] and st["done"] > 10000, st
        rng = np.random.default_rng(4)
        sample = np.sort(rng.choice(done, size=min(2000, len(done)), replace=False))
        assert np.array_equal(res[sample.astype(np.int64)], expected(spec, sample))
        h.trim()
        assert NatHeap(h).canonical_problems() == []
        for j in sample[[0, len(sample) // 2, len(sample) - 1]]:
            st1, r1 = kma_amd.serve_run(h, spec, jobs=[int(j)])
            assert not st1["aborted"] and st1["done"] == 1 and r1[0] == res[int(j)], (j, r1, res[int(j)])
        h.trim()
        assert NatHeap(h).canonical_problems() == []
