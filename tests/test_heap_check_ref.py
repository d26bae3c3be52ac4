"""Heap walker on the reference layout (kma_image_check) on the CPU.

Images come from the oracle's serialized replay (oracle/refmodel.c): a
fresh heap, a heap with live blocks in several classes (attached class
heads and detached superblocks) and the heap after every block was freed.
All must check clean, with the walk's counts equal to what the oracle
holds.  Then one mistake at a time is planted in a copy of the mid-trace
image, and each must be reported: the mistakes a broken allocator would
leave in this layout (a cycle in the free queue, a tail that is not the
last node, a queued superblock with a live bit, a free count that
disagrees with the bitmap, a class head left POISON or naming a queued
superblock, a bit past a superblock's slots, an orphan).  The reference's
queue validator (tb_clIndexedQueue.c:131-168) and its commented-out orphan
check (kma.cl:528-548) print and assert nothing, so the planted cases are
this walker's known answers."""
import numpy as np
import pytest

import kma_amd
from oracle import refmodel as rm
from tests import refimg

LOG2, S, BASE = 12, 24, 0x7f0000000000
HDR = 32 + 8 * (LOG2 - 1)
SB = 1 << LOG2


def check(img):
    r = kma_amd.HeapReport()
    rc = kma_amd.lib().kma_image_check(np.ascontiguousarray(img).ctypes.data, len(img), r)
    assert rc == 0, kma_amd.last_error()
    return r.as_dict()


@pytest.fixture(scope="module")
def mid():
    """Oracle heap after a random trace that leaves live blocks of several
    classes; returns (image, live block count, queue nodes)."""
    o = rm.RefHeap(LOG2, S, base=BASE)
    rng = np.random.default_rng(5)
    ops, live = refimg.random_trace(rng, 400, [8, 16, 64, 100, 500, 1020], free_frac=0.4, max_live=60)
    out, _ = o.replay(ops)
    assert (out[[j for j in live]] != 0).all()
    return o.image(), len(live), o.queue_walk(), o


def u32(img, off):
    return int(img[off:off + 4].view(np.uint32)[0])


def put32(img, off, v):
    img[off:off + 4] = np.array([v], dtype=np.uint32).view(np.uint8)


def put64(img, off, v):
    img[off:off + 8] = np.array([v], dtype=np.uint64).view(np.uint8)


def sb_off(s):
    return HDR + s * SB


def queue_nodes(img):
    """0-based superblock indexes of the free queue, head first."""
    i, out = (u32(img, 24) & 0x1fffe) >> 1, []
    while i:
        out.append(i - 1)
        i = (u32(img, sb_off(i - 1)) & 0x1fffe) >> 1
    return out


def class_heads(img):
    hd = img[32:HDR].view(np.uint64)
    return {c: (int(v) - (BASE + HDR)) // SB for c, v in enumerate(hd) if v > 1}


def test_fresh_heap_is_clean():
    o = rm.RefHeap(LOG2, S, base=BASE)
    r = check(o.image())
    assert r["problems"] == 0, r["first"]
    assert r["sb_free"] == S - 1 and r["sb_unqueued"] == 1 and r["live_blocks"] == 0 and r["sb_total"] == S


def test_mid_trace_and_free_all_are_clean(mid):
    img, nlive, qn, o = mid
    r = check(img)
    assert r["problems"] == 0, r["first"]
    assert r["live_blocks"] == nlive
    assert r["sb_free"] == qn
    assert r["sb_heads"] >= 3 and r["orphans"] == 0
    assert r["sb_heads"] + r["sb_detached"] + r["sb_free"] + r["sb_unqueued"] == S


def test_after_free_all_is_clean():
    o = rm.RefHeap(LOG2, S, base=BASE)
    rng = np.random.default_rng(8)
    ops, live = refimg.random_trace(rng, 500, [8, 16, 64, 100, 500], free_frac=0.4, max_live=50)
    o.replay(refimg.free_all(ops, live))
    r = check(o.image())
    assert r["problems"] == 0, r["first"]
    assert r["live_blocks"] == 0 and r["sb_heads"] == 0 and r["sb_detached"] == 0


def _in_use(img):
    """A superblock in use with at least two live blocks."""
    q = set(queue_nodes(img))
    for s in range(S - 1):
        if s in q:
            continue
        st = u32(img, sb_off(s) + 4)
        slots, fr = st >> 16, st & 0xffff
        if slots and slots - fr >= 2:
            return s, slots, fr
    raise AssertionError("no superblock in use")


def _plant_cycle(img):
    q = queue_nodes(img)
    put32(img, sb_off(q[-1]), (u32(img, sb_off(q[-1])) & ~0x1fffe) | (((q[1] + 1) << 1) & 0x1fffe))
    return "cycle"


def _plant_tail(img):
    q = queue_nodes(img)
    put32(img, 28, (u32(img, 28) & ~0x1fffe) | (((q[1] + 1) << 1) & 0x1fffe))
    return "tail names"


def _plant_queued_bit(img):
    q = queue_nodes(img)
    s = q[len(q) // 2]
    put32(img, sb_off(s) + SB - 4, 1)
    return "queued superblock %d" % s


def _plant_count(img):
    s, slots, fr = _in_use(img)
    put32(img, sb_off(s) + 4, (slots << 16) | (fr + 1))
    return "superblock %d counts" % s


def _plant_poison(img):
    put64(img, 32 + 8 * 7, 1)
    return "POISON"


def _plant_head_queued(img):
    q = queue_nodes(img)
    c = min(set(range(LOG2 - 1)) - set(class_heads(img)))
    put64(img, 32 + 8 * c, BASE + HDR + q[2] * SB)
    return "also the head of class %d" % c


def _plant_past_slots(img):
    s, slots, fr = _in_use(img)
    if slots % 32 == 0:
        pytest.skip("slots fill their last word")
    w = sb_off(s) + SB - 4 * (slots // 32 + 1)
    put32(img, w, u32(img, w) | (1 << (slots % 32)))
    return "past its %d slots" % slots


def _plant_orphan(img):
    heads = class_heads(img)
    c, s = sorted(heads.items())[0]
    st = u32(img, sb_off(s) + 4)
    slots = st >> 16
    put64(img, 32 + 8 * c, 0)                      # head dropped
    put32(img, sb_off(s) + 4, (slots << 16) | slots)  # every slot free
    for k in range((slots + 31) // 32):
        put32(img, sb_off(s) + SB - 4 * (k + 1), 0)
    return "orphan"


@pytest.mark.parametrize("plant", [_plant_cycle, _plant_tail, _plant_queued_bit, _plant_count, _plant_poison,
                                   _plant_head_queued, _plant_past_slots, _plant_orphan])
def test_planted_mistake_is_reported(mid, plant):
    img = mid[0].copy()
    want = plant(img)
    r = check(img)
    assert r["problems"] >= 1, plant.__name__
    assert any(want in line for line in r["first"]), (want, r["first"])


def test_not_a_heap_image():
    img = np.zeros(8192, dtype=np.uint8)
    r = kma_amd.HeapReport()
    assert kma_amd.lib().kma_image_check(img.ctypes.data, len(img), r) == -1
