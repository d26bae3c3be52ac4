"""The device-resident serving loop (kma_serve_run) on the GPU: continuous
admission (by projected use, or by a use watermark), one block per live job
per step, completed jobs' blocks freed.  Each job's result is the sum of the
patterns its blocks carried, so it must equal the closed form
sum_k pattern(j, k) and be the same when the job runs alone as in the mix
(tolerance: exact).  Under projected admission the use the loop projected
for every step without stalls must equal the blocks it held after that
step's growth phase (successful mallocs minus earlier frees)."""
import numpy as np
import pytest

import kma_amd
from tests.natimg import NatHeap

pytestmark = pytest.mark.gpu


def serve_heap(sblocks=256):
    # 1 MiB superblocks of 64 x 16 KiB blocks
    return kma_amd.Heap(layout=kma_amd.LAYOUT_NATIVE, sb_log2=20, sblocks=sblocks, classes=[16384])


def expected(spec, jobs):
    ln = kma_amd.serve_lengths(spec, jobs)
    out = np.zeros(len(jobs), dtype=np.uint64)
    for i, (j, L) in enumerate(zip(jobs, ln)):
        with np.errstate(over="ignore"):
            out[i] = kma_amd.serve_pattern(int(j), np.arange(int(L), dtype=np.uint64)).sum(dtype=np.uint64)
    return out


def check_stats(st):
    assert not st["aborted"]
    assert st["bad"] == 0
    assert st["mallocs"] == st["frees"] + st["drain_frees"], st
    assert st["admitted"] == st["done"] + st["drained"] + st["preempted"], st
    if st["admit"] == kma_amd.ADMIT_PROJECTED:
        assert st["proj_mismatch"] == 0, st
        assert st["proj_checked"] + st["stalls"] > 0 or st["steps"] <= 1, st


def test_serve_mix_and_alone():
    print("serve mix", flush=True)
    spec = kma_amd.serve_spec(rate_per_ms=40.0, len_mu=3.5, len_sigma=1.0, duration_ms=120.0, max_len=1024,
                              max_live=2048, nwg=64, seed=4)
    with serve_heap() as h:
        st, res = kma_amd.serve_run(h, spec, nresults=1 << 20)
        check_stats(st)
        assert st["done"] > 100, st
        done = np.nonzero(res)[0].astype(np.uint64)
        assert len(done) == st["done"]
        assert np.array_equal(res[done.astype(np.int64)], expected(spec, done))
        h.trim()
        assert NatHeap(h).canonical_problems() == []
        # the same jobs alone
        pick = done[[0, len(done) // 2, len(done) - 1]]
        for j in pick:
            st1, r1 = kma_amd.serve_run(h, spec, jobs=[int(j)])
            check_stats(st1)
            assert st1["done"] == 1 and r1[0] == res[int(j)], (j, r1, res[int(j)])
        h.trim()
        assert NatHeap(h).canonical_problems() == []


def test_serve_memory_pressure():
    """Watermark admission on a heap far smaller than the demand: admission
    stops at the watermark, jobs stall on NULL and are preempted; results
    stay exact."""
    print("serve pressure, watermark admission", flush=True)
    spec = kma_amd.serve_spec(rate_per_ms=200.0, len_mu=4.0, len_sigma=1.0, duration_ms=80.0, max_len=1024,
                              max_live=4096, nwg=64, watermark_pct=80, seed=5, admit=kma_amd.ADMIT_WATERMARK)
    with serve_heap(sblocks=16) as h:  # 16 MiB: 1,024 blocks
        st, res = kma_amd.serve_run(h, spec, nresults=1 << 20)
        check_stats(st)
        assert st["admitted"] - st["preempted"] < st["jobs"], st  # admission held jobs back
        assert st["stalls"] > 0 and st["preempted"] > 0, st  # the heap ran out and jobs were preempted
        done = np.nonzero(res)[0].astype(np.uint64)
        assert len(done) > 0
        assert np.array_equal(res[done.astype(np.int64)], expected(spec, done))
        h.trim()
        assert NatHeap(h).canonical_problems() == []


def test_serve_projected_pressure():
    """Projected admission on a heap far smaller than the demand: jobs wait
    for room instead of stalling (a stall is left only where an XCD's own
    superblocks run short, which on a heap this small one long job can
    cause); the projection of every step without stalls equals the blocks
    held; results exact."""
    print("serve pressure, projected admission", flush=True)
    spec = kma_amd.serve_spec(rate_per_ms=200.0, len_mu=4.0, len_sigma=1.0, duration_ms=80.0, max_len=1024,
                              max_live=4096, nwg=64, watermark_pct=75, seed=5)
    with serve_heap(sblocks=1024) as h:  # 1 GiB: 65,536 blocks, a few % of the demand
        st, res = kma_amd.serve_run(h, spec, nresults=1 << 20)
        print("projected pressure: %s" % {k: v for k, v in st.items() if not isinstance(v, float)}, flush=True)
        check_stats(st)
        assert st["cap_blocks"] == 1024 * 64 * 75 // 100, st
        assert st["admitted"] < st["jobs"], st  # admission held jobs back
        assert st["peak_proj_frac"] > 0.7, st  # and the heap was used up to the limit
        assert st["stalls"] <= 0.01 * st["mallocs"] and st["preempted"] <= 0.05 * st["admitted"], st
        assert st["proj_checked"] >= 0.9 * st["steps"], st
        done = np.nonzero(res)[0].astype(np.uint64)
        assert len(done) == st["done"] > 0
        assert np.array_equal(res[done.astype(np.int64)], expected(spec, done))
        h.trim()
        assert NatHeap(h).canonical_problems() == []


def test_serve_projected_overcommit():
    """Projected admission told to admit 150 % of the heap: the heap runs
    out, jobs stall, the youngest are preempted and the timeline is rebuilt
    from the live slots; results stay exact and the heap comes back."""
    print("serve overcommit, projected admission", flush=True)
    spec = kma_amd.serve_spec(rate_per_ms=200.0, len_mu=4.0, len_sigma=1.0, duration_ms=80.0, max_len=1024,
                              max_live=4096, nwg=64, watermark_pct=150, seed=6)
    with serve_heap(sblocks=16) as h:
        st, res = kma_amd.serve_run(h, spec, nresults=1 << 20)
        print("projected overcommit: %s" % {k: v for k, v in st.items() if not isinstance(v, float)}, flush=True)
        check_stats(st)
        assert st["stalls"] > 0 and st["preempted"] > 0 and st["rebuilds"] > 0, st
        done = np.nonzero(res)[0].astype(np.uint64)
        assert len(done) == st["done"] > 0
        assert np.array_equal(res[done.astype(np.int64)], expected(spec, done))
        h.trim()
        assert NatHeap(h).canonical_problems() == []


def test_serve_spec_checks():
    """Projected admission looks at most KMA_SERVE_MAX_PROJECTED_LEN steps
    ahead; the watermark policy takes at most 100 %."""
    with serve_heap(sblocks=16) as h:
        with pytest.raises(kma_amd.KmaError):
            kma_amd.serve_run(h, kma_amd.serve_spec(duration_ms=1.0, max_len=8193, max_live=256))
        with pytest.raises(kma_amd.KmaError):
            kma_amd.serve_run(h, kma_amd.serve_spec(duration_ms=1.0, max_len=64, max_live=256, watermark_pct=101,
                                                    admit=kma_amd.ADMIT_WATERMARK))
        with pytest.raises(kma_amd.KmaError):
            kma_amd.serve_run(h, kma_amd.serve_spec(duration_ms=1.0, max_len=64, max_live=256, admit=2))
