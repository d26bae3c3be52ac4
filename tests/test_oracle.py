"""CPU tests that pin the oracle (oracle/refmodel.c) before anything is
compared against it.

The reference ships no golden vectors (its tests print and never assert),
so the oracle is pinned by: values derived by hand from the reference source,
a closed-form derivation of the heap-init image, brute-force properties of
the size-class functions, and invariants of serialized traces.
"""
import numpy as np
import pytest

from oracle import refmodel as rm
from tests import refimg


# --- size classes: tables derived by hand from kma.cl:55-114 ----------------
# 4 KiB: classes 0..4 are 4..64 B (2^(c+2) while c+2 <= 6); from class 5 the
# size is floor((4096 - 12 - bitmap bytes) / items) & ~3.
TABLE_4K = [4, 8, 16, 32, 64, 124, 252, 508, 1020, 2040, 4080]
SLOTS_4K = [989, 502, 253, 127, 63, 32, 16, 8, 4, 2, 1]
TABLE_16K = [4, 8, 16, 32, 64, 128, 252, 508, 1020, 2044, 4092, 8184, 16368]
SLOTS_16K = [3965, 2014, 1015, 509, 255, 127, 64, 32, 16, 8, 4, 2, 1]


def test_size_tables_match_hand_derivation():
    assert [rm.size_by_sbid(12, c) for c in range(11)] == TABLE_4K
    assert [rm.slots_by_size(12, s) for s in TABLE_4K] == SLOTS_4K
    assert [rm.size_by_sbid(14, c) for c in range(13)] == TABLE_16K
    assert [rm.slots_by_size(14, s) for s in TABLE_16K] == SLOTS_16K
    assert rm.size_by_sbid(12, -1) == 0 and rm.size_by_sbid(12, 12) == 0


@pytest.mark.parametrize("size,sbid", [(4, 0), (5, 1), (64, 4), (65, 5), (124, 5), (125, 6),
                                        (128, 6), (4080, 10), (4081, -1), (0, -1)])
def test_sbid_known_answers_4k(size, sbid):
    assert rm.sbid_by_size(12, size) == sbid


def test_sbid_known_answers_16k():
    assert rm.sbid_by_size(14, 128) == 5
    assert rm.sbid_by_size(14, 129) == 6
    assert rm.sbid_by_size(14, 16368) == 12
    assert rm.sbid_by_size(14, 16369) == -1
    assert rm.slots_by_size(14, 8) == 2014


@pytest.mark.parametrize("log2", [12, 14])
def test_sbid_is_smallest_fitting_class_bruteforce(log2):
    """For every size 4..SB-16 the class is the smallest one whose block
    holds the request (brute force over the whole range)."""
    nb = log2 - 1
    table = [rm.size_by_sbid(log2, c) for c in range(nb)]
    for size in range(4, (1 << log2) - 16 + 1):
        c = rm.sbid_by_size(log2, size)
        want = next(i for i, t in enumerate(table) if t >= size)
        assert c == want, size


@pytest.mark.parametrize("log2", [12, 14])
def test_blocks_and_bitmap_fit_in_superblock(log2):
    """Layout invariant: 12-byte header + slots blocks + ceil(slots/32) bitmap
    words never exceed the superblock (the bitmap grows down from the end)."""
    sb = 1 << log2
    for c in range(log2 - 1):
        size = rm.size_by_sbid(log2, c)
        slots = rm.slots_by_size(log2, size)
        assert slots >= 1
        assert 12 + slots * size + 4 * ((slots + 31) // 32) <= sb, (c, size, slots)


# --- heap image ---------------------------------------------------------------
def test_header_sizes():
    # sizeof(struct clheap) on a 64-bit device: 8 + 24 + 8 * buckets
    assert rm.header_bytes(12) == 120
    assert rm.header_bytes(14) == 136


@pytest.mark.parametrize("log2,S", [(12, 2), (12, 3), (12, 4), (12, 127), (12, 4096), (14, 4096), (12, 2100)])
def test_init_image_equals_closed_form(log2, S):
    base = 0x7F0000000000
    h = rm.RefHeap(log2, S, base=base)
    want = refimg.closed_form_init(log2, S, base)
    assert np.array_equal(h.image(), want)
    assert h.queue_walk() == S - 1


def test_first_allocations_known_answers():
    # fresh 4 KiB heap: sb[0] (offset 120) is popped first; a fresh
    # superblock hands out its top slot first (kma.cl:180,192)
    h = rm.RefHeap(12, 4096)
    assert h.malloc(64) - h.base == 120 + 12 + 62 * 64 == 4100
    assert h.malloc(64) - h.base == 4036
    h2 = rm.RefHeap(12, 4096)
    assert h2.malloc(8) - h2.base == 120 + 12 + 501 * 8 == 4140
    h3 = rm.RefHeap(14, 4096)
    assert h3.malloc(64) - h3.base == 136 + 12 + 254 * 64 == 16404


def test_malloc_free_exact_bytes_hand_derived():
    """4 KiB superblocks, 4 of them: malloc(64) then free.  Every byte below
    was derived by hand from kma.cl:150-373 and clIndexedQueue.cl:119-209."""
    base = 0x1000000
    h = rm.RefHeap(12, 4, base=base)
    p = h.malloc(64)
    assert p - base == 4100
    h.free(p)
    img = h.image()
    hdr = 120
    assert refimg.u32(img, 24) == 4 | (2 << 22)      # head: sb1, tag 2
    assert refimg.u32(img, 28) == 2 | (4 << 22)      # tail: sb0, tag 4
    assert refimg.u32(img, hdr + 0) == 2 << 22        # sb0.next: NULL, tag 2
    assert refimg.u32(img, hdr + 4) == 63 << 16       # state: 63 slots, queued
    assert refimg.u32(img, hdr + 8) == 64             # block size
    assert refimg.u32(img, hdr + 12 + 62 * 64) == 0x40000000  # old bitmap word
    assert refimg.u32(img, hdr + 4096 - 8) == 0       # bitmap word 1 cleared
    assert refimg.u32(img, hdr + 2 * 4096) == 2 | (1 << 22)   # sb2.next -> sb0
    assert refimg.u64(img, 32 + 8 * 4) == 0           # class head 4 NULL
    assert h.queue_walk() == 3


def test_fifo_superblock_order_and_exhaustion():
    # S = 8: sb[0..6] are queued, the last queued node always stays behind as
    # the dummy, so S-2 = 6 superblocks of 2 slots (2040-B class) are usable
    h = rm.RefHeap(12, 8)
    offs = [h.malloc(2000) - h.base for _ in range(12)]
    sbs = [(o - 120) // 4096 for o in offs]
    assert sbs == [0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5]
    assert h.malloc(2000) == 0
    assert h.malloc(8) == 0


def test_one_slot_class_reports_hang():
    h = rm.RefHeap(12, 16)
    assert h.malloc(4080) != 0
    with pytest.raises(RuntimeError):
        h.malloc(4080)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_trace_invariants(seed):
    """Serialized random traces: every live block in-heap, inside its
    superblock's data area, class-correct and non-overlapping; after freeing
    everything the heap is in canonical free form."""
    rng = np.random.default_rng(seed)
    log2, S = 12, 64
    sizes = [1, 3, 4, 7, 8, 15, 16, 33, 64, 100, 124, 200, 252, 500, 1020, 2040]
    ops, live = refimg.random_trace(rng, 3000, sizes, max_live=400)
    h = rm.RefHeap(log2, S)
    out, nulls = h.replay(ops)
    hdr = 120
    spans = []
    for j in live:
        p = int(out[j])
        if not p:
            continue
        off = p - h.base
        sb = (off - hdr) // 4096
        assert 0 <= sb < S - 1
        size = rm.size_by_sbid(12, rm.sbid_by_size(12, int(ops[j])))
        assert size >= ops[j]
        inner = off - hdr - sb * 4096
        assert inner >= 12 and (inner - 12) % size == 0
        slots = rm.slots_by_size(12, size)
        assert (inner - 12) // size < slots
        spans.append((off, off + size))
    spans.sort()
    for a, b in zip(spans, spans[1:]):
        assert a[1] <= b[0]
    h2 = rm.RefHeap(log2, S)
    h2.replay(refimg.free_all(ops, live))
    assert refimg.ref_canonical_free(h2.image(), log2, S) == []


def test_mix_is_splitmix_of_linear_combination():
    # splitmix64 finaliser with the published constants (Steele et al.)
    def sm(z):
        M = (1 << 64) - 1
        z = (z + 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    M = (1 << 64) - 1
    for s, g, r, k in [(1, 0, 0, 0), (3, 524287, 49, 1), (2, 12345, 7, 0)]:
        lin = (s * 0x9E3779B97F4A7C15 + g * 0xBF58476D1CE4E5B9 + r * 0x94D049BB133111EB
               + k * 0xD6E8FEB86659FD93) & M
        assert rm.mix(s, g, r, k) == sm(lin)
    # seed 0, everything 0: splitmix64 of 0 is the first output of the
    # published generator seeded with 0
    assert rm.mix(0, 0, 0, 0) == 0xE220A8397B1DCDAF


def test_cpu_restatement_plain_small():
    # The multi-threaded restatement keeps the reference's free(), which
    # stores the old bitmap word into the block after releasing its bit
    # (kma.cl:338): a block re-allocated in that window can have its first
    # word overwritten, so a few read-back mismatches are the reference's own.
    r = rm.cpu_run(12, 4096, 4, 1024, 100, 0, 64)
    assert r["ops"] == 204800 and r["nulls"] == 0
    assert r["bad"] < 204800 // 100
