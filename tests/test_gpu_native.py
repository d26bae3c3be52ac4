"""GPU tests of the native gfx950 layout (kma_device.h) through libkma.so.

The native layout has no counterpart in the reference, so it is checked by
invariants: every block in-heap, size-class-correct, aligned and
non-overlapping; seeded patterns read back; and after a full malloc/free
trace the metadata returns to its initial image apart from generation and
ABA-tag fields and pool order."""
import numpy as np
import pytest

import kma_amd
from tests import natimg, refimg

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_init_is_canonical():
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 1024, npages=256) as h:
        nh = natimg.NatHeap(h)
        assert nh.classes == [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192]
        assert nh.caps == [(1 << 14) // c for c in nh.classes]
        assert nh.off_data % (1 << 14) == (-h.device_ptr) % (1 << 14)
        assert nh.canonical_problems() == []


@pytest.mark.parametrize("seed", range(3))
def test_serial_trace_invariants(seed):
    rng = np.random.default_rng(300 + seed)
    sizes = [1, 7, 8, 9, 16, 24, 64, 100, 128, 500, 512, 1000, 2048, 4000, 8192, 8193,
             16384, 20000, 65536]
    ops, live = refimg.random_trace(rng, 4000, sizes, max_live=600)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 256, nctx=8, npages=512) as h:
        _say("native serial invariants seed=%d" % seed)
        out = h.trace_serial(ops)
        nh = natimg.NatHeap(h)
        offs = [int(out[j]) for j in live]
        sizes_live = [int(ops[j]) for j in live]
        assert sum(1 for o, s in zip(offs, sizes_live) if o == 0 and s > 0) == 0
        assert natimg.check_blocks(nh, offs, sizes_live) == []
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 256, nctx=8, npages=512) as h:
        h.trace_serial(refimg.free_all(ops, live))
        h.trim()
        nh = natimg.NatHeap(h)
        assert nh.canonical_problems() == []
        assert int(nh.meta()["stats"][0]) == 0  # no double free


def _fill_and_free(h, size, n_try):
    """One serialized trace: n_try mallocs of `size`, then a free of every
    success; returns the number of successes."""
    out = h.trace_serial([size] * n_try + [-(i + 1) for i in range(n_try)])
    return int(np.count_nonzero(out[:n_try]))


def test_page_region_fills_exactly():
    """A 20,000-page region (five summary words per window level, so the
    top-level summary decides which summary words are read): runs of one
    size fill it exactly, i.e. every aligned window is found and none is
    handed out twice (1-page runs: every page; 2 and 4: every aligned pair
    and quad, the 32-page last word included; 64: the 312 full words), and
    it comes back to the empty image."""
    P = 20000
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 16, nctx=8, npages=P) as h:
        for pages, expect in ((1, P), (2, P // 2), (4, P // 4), (64, P // 64), (1, P)):
            _say("page fill: %d-page runs" % pages)
            assert _fill_and_free(h, pages << 14, expect + 1) == expect, pages
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


def test_serial_size_zero_and_oversize():
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 64, nctx=8, npages=64) as h:
        out = h.trace_serial([0, 8, -2, -1, (64 << 14) + 1, 1 << 20, -6])
        assert out[0] == 0 and out[1] != 0 and out[4] == 0
        assert out[5] != 0  # 64 pages of 16 KiB: the whole page region
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


def test_double_free_detected_not_applied():
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 64, nctx=8) as h:
        out = h.trace_serial([64, 64, -1, -1, -2])
        assert out[0] and out[1]
        nh = natimg.NatHeap(h)
        assert int(nh.meta()["stats"][0]) == 1
        h.trim()
        assert nh.canonical_problems() == []


def test_free_into_released_superblock_is_a_bad_free():
    """8 KiB blocks, 2 per 16 KiB superblock: the third malloc finds the
    first superblock full and detaches it; freeing both of its blocks
    releases it to the pool (no class any more); freeing one of them again
    is then counted as a bad free, not a double free, and changes nothing."""
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 16, nctx=8) as h:
        out = h.trace_serial([8192, 8192, 8192, -1, -2, -1, -3])
        assert out[0] and out[1] and out[2]
        assert out[0] >> 14 == out[1] >> 14 != out[2] >> 14
        st = natimg.NatHeap(h).meta()["stats"]
        assert int(st[1]) == 1 and int(st[0]) == 0, list(st)
        r = h.check()
        assert r["problems"] == 0 and r["live_blocks"] == 0, r
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


def test_wave_free_runs_and_double_frees():
    """One wavefront frees 64 blocks of consecutive slots: the lanes' bits go
    back in runs, one RMW per run and bitmap word.  A lane that frees its
    neighbour's block (a double free inside a run), or a block freed before
    (at the start of a run), is counted as a double free and changes
    nothing; every other block is freed, and the heap ends canonical."""
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 64, nctx=8) as h:
        stats = lambda: [int(x) for x in natimg.NatHeap(h).meta()["stats"][:2]]
        offs = h.trace_serial([64] * 64).copy()
        assert np.all(offs) and np.all(np.diff(offs.astype(np.int64)) == 64)
        wf = offs.copy()
        wf[20] = offs[19]                  # lane 20 frees lane 19's block again
        h.wave_free(wf)
        assert stats() == [1, 0]
        h.wave_free([offs[20]])            # the block lane 20 left live
        assert stats() == [1, 0]
        r = h.check()
        assert r["problems"] == 0 and r["live_blocks"] == 0, r
        again = h.trace_serial([64] * 64).copy()
        h.wave_free([again[32]])
        h.wave_free(again)                 # lane 32 starts a run with a freed block
        assert stats() == [2, 0]
        r = h.check()
        assert r["problems"] == 0 and r["live_blocks"] == 0, r
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


def test_exhaustion_returns_null_then_recovers():
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 8, nctx=8, npool=16) as h:
        # 8 superblocks of 2 x 8 KiB blocks
        ops = [8192] * 20
        frees = [-(i + 1) for i in range(20)]
        out = h.trace_serial(ops + frees + [8192] * 16)
        assert int(np.count_nonzero(out[:20])) == 16
        assert int(np.count_nonzero(out[40:])) == 16


def test_plain_small_native():
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 12, 4096) as h:
        _say("native plain small")
        st = h.run(kma_amd.WL_FIXED, 1024, 100, block=64, fixed=64)
        assert st["ops"] == 204800 and st["nulls"] == 0 and st["bad"] == 0
        assert st["double_frees"] == 0 and st["bad_frees"] == 0
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


@pytest.mark.parametrize("form", [kma_amd.FORM_PLAIN, kma_amd.FORM_FUSED, kma_amd.FORM_BATCHED])
def test_full_occupancy_native(form):
    """Full-occupancy workload (1 GiB heap, 16 KiB superblocks, 524,288
    work-items, log-uniform 8 B..2 KiB, 0..64 holds), 3 rounds."""
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 65536) as h:
        _say("native full occupancy form=%d" % form)
        st = h.run(kma_amd.WL_LOGUNI, 524288, 3, form=form, block=256, seed=1)
        assert st["ops"] == 2 * 3 * 524288, st
        assert st["nulls"] == 0 and st["bad"] == 0 and st["waits"] == [0, 0, 0, 0], st
        assert st["double_frees"] == 0 and st["bad_frees"] == 0
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


@pytest.mark.parametrize("mode", [kma_amd.WL_REF_TEST, kma_amd.WL_LOWVAR, kma_amd.WL_HIGHVAR])
def test_reference_kernels_full_occupancy_native(mode):
    """The reference's bench kernels (kma.cl:424-526) at 524,288 work-items
    on a native heap, one launch of 20 iterations: every lane of a wavefront
    asks for the same one or two classes, so most of each wavefront's
    reserved slots are taken by scattered live blocks and its lanes scan for
    others.  Exact op counts, no NULL, no watchdog expiry, and a time bound
    that the lanes' scans colliding on the same free bits would break
    (0.377 s for 100 iterations of kma_test_malloc when they did, 0.026 s
    without)."""
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 16384) as h:
        _say("reference kernel mode=%d, full occupancy, native" % mode)
        st = h.run(mode, 524288, 20, form=kma_amd.FORM_FUSED, block=256)
        extra = 0 if mode == kma_amd.WL_REF_TEST else 2 * 20
        assert st["ops"] == 2 * 524288 * 20 + extra, st
        assert st["nulls"] == 0 and st["bad"] == 0 and st["waits"] == [0, 0, 0, 0], st
        assert st["double_frees"] == 0 and st["bad_frees"] == 0, st
        assert st["gpu_ms"] < 40, st
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


@pytest.mark.parametrize("no_partial", [1, 0])
@pytest.mark.parametrize("form", [kma_amd.FORM_PLAIN, kma_amd.FORM_BATCHED])
def test_cross_xcd_supply_by_migration(form, no_partial):
    """A heap of 16 superblocks (two per XCD, four 4 KiB blocks each) under
    eight work-groups of 4 KiB requests: superblocks emptied on one XCD stay
    in its dirty shards, so the others get them only through in-kernel
    migrations (dirty -> clean after one L2 write-back).  Patterns read back,
    no watchdog fires, and the heap is canonical after trim.  The per-lane
    form requests migrations (its superblocks idle in dirty shards while the
    owner runs); the collective form takes whole superblocks and mostly
    re-pops its own dirty ones, so migrations there are timing-dependent.
    With partial reuse on (no_partial 0) an XCD under pool pressure also
    lists and refills from its own partly free superblocks (four slots
    each here), so a run may need no migration at all: that case checks
    only correctness."""
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 16, nctx=8, npool=16, no_partial=no_partial) as h:
        _say("native cross-XCD supply form=%d no_partial=%d" % (form, no_partial))
        st = h.run(kma_amd.WL_FIXED, 8 * 256, 30, form=form, block=256, fixed=4096)
        assert st["ops"] >= 30 * 8 * 256 and st["bad"] == 0, st
        assert st["double_frees"] == 0 and st["bad_frees"] == 0, st
        assert st["waits"] == [0, 0, 0, 0], st
        assert 0 < st["nulls"] < st["mallocs"], st
        if form == kma_amd.FORM_PLAIN and no_partial:
            assert st["migrated"] > 0, st
        if no_partial:
            assert st["listed"] == 0 and st["reattached"] == 0, st
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


def test_batched_plain_small_and_partial_workgroup():
    """Batched form with a work-item count that leaves the last work-group
    partly empty, and the fixed-size workload."""
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 12, 4096) as h:
        _say("native batched fixed 1000 wi")
        st = h.run(kma_amd.WL_FIXED, 1000, 50, form=kma_amd.FORM_BATCHED, block=256, fixed=64)
        assert st["ops"] == 2 * 50 * 1000 and st["nulls"] == 0 and st["bad"] == 0, st
        _say("native batched loguni 777 wi")
        st = h.run(kma_amd.WL_LOGUNI, 777, 20, form=kma_amd.FORM_BATCHED, block=256, seed=5)
        assert st["ops"] == 2 * 20 * 777 and st["nulls"] == 0 and st["bad"] == 0, st
        assert st["double_frees"] == 0 and st["bad_frees"] == 0
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


def test_batched_rejects_reference_layout():
    with kma_amd.Heap(kma_amd.LAYOUT_REFERENCE, 12, 64) as h:
        with pytest.raises(kma_amd.KmaError):
            h.run(kma_amd.WL_FIXED, 64, 1, form=kma_amd.FORM_BATCHED, block=256)


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_batched_mixed_with_per_lane_api(variant):
    """Blocks from the work-group collective malloc freed by the per-lane
    kma_free (variant 2), and the reverse (variant 1): both forms share one
    heap protocol."""
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 65536) as h:
        _say("native batched variant=%d" % variant)
        st = h.run(kma_amd.WL_LOGUNI, 65536, 3, form=kma_amd.FORM_BATCHED, block=256, seed=2,
                   variant=variant)
        assert st["ops"] == 2 * 3 * 65536 and st["nulls"] == 0 and st["bad"] == 0, st
        assert st["double_frees"] == 0 and st["bad_frees"] == 0 and st["waits"] == [0, 0, 0, 0], st
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []


def _expected_claim(bm, cap, base, got, d0, size):
    """Model of the group claim with no concurrent claimer: first the free
    bits among the reserved slots [base, base + got) (clipped to the
    capacity), then words from (base/32) mod W onward, lowest free bits
    first; ranks in that order."""
    W = (cap + 31) // 32
    w0 = (base >> 5) % W
    out, bm = [], list(bm)
    for b in range(base, min(base + got, cap)):
        if not (bm[b >> 5] >> (b & 31)) & 1:
            bm[b >> 5] |= 1 << (b & 31)
            out.append(d0 + b * size)
    for k in range(W):
        wi = (w0 + k) % W
        valid = cap - wi * 32
        for b in range(min(32, valid)):
            if len(out) == got:
                break
            if not (bm[wi] >> b) & 1:
                bm[wi] |= 1 << b
                out.append(d0 + (wi * 32 + b) * size)
    return out, bm


def test_wave_claim_matches_model():
    rng = np.random.default_rng(11)
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 256, nctx=8) as h:
        nh = natimg.NatHeap(h)
        W = nh.bm_words
        cls, base, got, bms = [], [], [], []
        for i in range(64):
            c = i % nh.nclasses
            cap = nh.caps[c]
            words = (cap + 31) // 32
            dens = rng.choice([0.0, 0.3, 0.7, 0.95])
            bits = (rng.random(words * 32) < dens)
            bits[cap:] = False
            bm = np.zeros(W, dtype=np.uint32)
            for wi in range(words):
                bm[wi] = int(sum(1 << b for b in range(32) if bits[wi * 32 + b]))
            nfree = cap - int(bits.sum())
            cls.append(c)
            base.append(int(rng.integers(0, cap)))
            got.append(int(min(nfree, rng.integers(0, 65))))
            bms.append(bm)
        res, bm_out, done = kma_amd.check_wave_claim(h, cls, base, got, np.concatenate(bms))
        for i in range(64):
            c = cls[i]
            d0 = nh.off_data + (i << nh.sb_log2)
            want, wbm = _expected_claim([int(x) for x in bms[i]], nh.caps[c], base[i], got[i], d0,
                                        nh.classes[c])
            assert done[i] == got[i], (i, c, done[i], got[i])
            assert [int(x) for x in res[i][:got[i]]] == want, (i, c, base[i], got[i])
            assert [int(x) for x in bm_out[i * W:(i + 1) * W]] == [int(x) for x in wbm[:W]], i


def test_batched_single_workgroup():
    """One work-group, no other allocator activity: the collective malloc and
    free on mixed classes."""
    with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, 14, 4096) as h:
        st = h.run(kma_amd.WL_LOGUNI, 256, 20, form=kma_amd.FORM_BATCHED, block=256, seed=3)
        assert st["ops"] == 2 * 20 * 256 and st["nulls"] == 0 and st["bad"] == 0, st
        h.trim()
        assert natimg.NatHeap(h).canonical_problems() == []
