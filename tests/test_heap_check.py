"""Heap walker / orphan detector (kma_image_check) on the CPU.

The CPU simulation (tests/cpu_sim) runs the batched workload on a native
heap in host memory and writes the metadata image after the run (context
heads attached, spares and dirty shards filled) and after the trim (every
superblock back in a clean shard).  Both must check clean; then one mistake
at a time is planted in a copy of the image, and each must be reported.
The mistakes planted are those a broken allocator would leave: a count that
disagrees with the bitmap, a superblock lost from every free list, one in
two free lists, an orphan, a stale shard count or mask bit, a page run
without its length, a wrong summary bit.  The reference's heap walker is
commented out (kma.cl:528-548), so this checker has no reference output to
match; the planted cases are its known answers."""
import os
import subprocess

import numpy as np
import pytest

import kma_amd
from tests import natimg

HERE = os.path.dirname(os.path.abspath(__file__))
SIM_DIR = os.path.join(HERE, "cpu_sim")
SIM = os.path.join(SIM_DIR, "_build", "kma_sim")


@pytest.fixture(scope="module")
def images(tmp_path_factory):
    subprocess.check_call(["make", "-s", "-C", SIM_DIR])
    d = tmp_path_factory.mktemp("img")
    path = str(d / "heap")
    # 16 KiB superblocks, 1024 of them, 4 work-groups x 3 rounds of the
    # full-occupancy size mix, 64 contexts, 16 shards, 256 pages
    env = dict(os.environ, KMA_SIM_DUMP=path)
    out = subprocess.run([SIM, "14", "1024", "4", "1024", "3", "1", "1", "64", "16", "0", "256"], env=env,
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return {s: np.fromfile(path + "." + s, dtype=np.uint8) for s in ("run", "trim")}


class Img:
    """Writable view of a metadata image's tables."""

    def __init__(self, img):
        self.b = img.copy()
        self.h = np.frombuffer(self.b[: natimg.HDR_DTYPE.itemsize].tobytes(), dtype=natimg.HDR_DTYPE)[0]
        g = lambda k: int(self.h[k])
        self.nsb, self.nctx, self.npool, self.W = g("nsb"), g("nctx"), g("npool"), g("bm_words")
        self.npages = g("npages")
        self.ns = ((self.npages + 63) // 64 + 63) // 64
        self.ns2 = (self.ns + 63) // 64

    def view(self, key, n, dt, extra=0):
        off = int(self.h[key]) + extra
        return self.b[off: off + n * np.dtype(dt).itemsize].view(dt)

    @property
    def state(self):
        return self.view("off_state", self.nsb, "<u8")

    @property
    def next(self):
        return self.view("off_next", self.nsb, "<u4")

    @property
    def bitmap(self):
        return self.view("off_bitmap", self.nsb * self.W, "<u4").reshape(self.nsb, self.W)

    @property
    def heads(self):
        return self.view("off_heads", self.nctx * natimg.HEADS_PER_CTX, "<u8").reshape(self.nctx, -1)

    @property
    def pool(self):
        return self.view("off_pool", self.npool * natimg.POOL_STRIDE, "<u8").reshape(self.npool, -1)

    @property
    def pmask(self):
        return self.view("off_stats", 8 * 16, "<u8", 256)[::16]

    @property
    def spare(self):
        return self.view("off_spare", self.nctx * 32, "<u4").reshape(self.nctx, 32)

    @property
    def pbits(self):
        return self.view("off_pbitmap", (self.npages + 63) // 64, "<u8")

    @property
    def plen(self):
        return self.view("off_plen", self.npages, "u1")

    @property
    def psum(self):
        return self.view("off_psum", natimg.PAGE_LEVELS * self.ns, "<u8").reshape(natimg.PAGE_LEVELS, -1)

    @property
    def psum2(self):
        return self.view("off_psum", natimg.PAGE_LEVELS * self.ns2, "<u8",
                         natimg.PAGE_LEVELS * self.ns * 8).reshape(natimg.PAGE_LEVELS, -1)

    def owned(self):
        """(ctx, class, sb) of every attached head."""
        out = []
        for c in range(self.nctx):
            for k in range(natimg.HEADS_PER_CTX):
                v = int(self.heads[c, k])
                if v > 1:
                    out.append((c, k, (v & 0xFFFFFFFF) - 2))
        return out

    def nonempty_shard(self):
        for s in range(self.npool):
            if int(self.pool[s, 0]) & 0xFFFFFFFF:
                return s
        raise AssertionError("no pooled superblock")

    def check(self):
        return kma_amd.image_check(self.b)


def _problems(r, *words):
    text = "\n".join(r["first"])
    assert r["problems"] > 0, r
    for w in words:
        assert w in text, text


def test_clean_images(images):
    run, trim = Img(images["run"]), Img(images["trim"])
    r = run.check()
    assert r["problems"] == 0, r["first"]
    assert r["sb_total"] == 1024 and r["orphans"] == 0 and r["live_blocks"] == 0
    assert r["sb_heads"] == len(run.owned()) > 0
    assert r["sb_free"] + r["sb_heads"] == 1024
    r = trim.check()
    assert r["problems"] == 0, r["first"]
    assert r["sb_free"] == 1024 and r["sb_heads"] == 0 and r["sb_in_spares"] == 0
    assert r["pages_total"] == 256 and r["page_runs"] == 0


def test_live_block_is_consistent(images):
    m = Img(images["run"])
    c, k, sb = m.owned()[0]
    m.bitmap[sb, 0] |= 1
    m.state[sb] += 1
    r = m.check()
    assert r["problems"] == 0, r["first"]
    assert r["live_blocks"] == 1 and r["live_bytes"] == int(m.h["class_size"][k])


def test_count_disagrees_with_bitmap(images):
    m = Img(images["run"])
    sb = m.owned()[0][2]
    m.bitmap[sb, 0] |= 1
    _problems(m.check(), "counts 0 blocks, 1 bits are set")


def test_bit_past_the_class_slots(images):
    m = Img(images["run"])
    for c, k, sb in m.owned():
        cap = int(m.h["class_cap"][k])
        if cap < 32 * m.W:
            m.bitmap[sb, cap // 32] |= np.uint32(1 << (cap % 32))
            m.state[sb] += 1
            _problems(m.check(), "bits past its %d slots" % cap)
            return
    pytest.skip("every attached class fills its bitmap")


def test_orphan(images):
    m = Img(images["run"])
    c, k, sb = m.owned()[0]
    m.heads[c, k] = 0
    m.state[sb] &= ~np.uint64(1 << 37)
    r = m.check()
    assert r["orphans"] == 1
    _problems(r, "superblock %d" % sb, "orphan")


def test_head_without_owned_bit(images):
    m = Img(images["run"])
    c, k, sb = m.owned()[0]
    m.state[sb] &= ~np.uint64(1 << 37)
    _problems(m.check(), "context %d head %d names superblock %d" % (c, k, sb))


def test_superblock_lost_from_its_shard(images):
    m = Img(images["trim"])
    s = m.nonempty_shard()
    top = (int(m.pool[s, 0]) & 0xFFFFFFFF) - 1
    m.pool[s, 0] = (int(m.pool[s, 0]) & ~0xFFFFFFFF) | int(m.next[top])
    _problems(m.check(), "free superblock %d is in 0 free lists" % top, "shard %d holds" % s)


def test_superblock_in_two_free_lists(images):
    m = Img(images["trim"])
    s = m.nonempty_shard()
    top = (int(m.pool[s, 0]) & 0xFFFFFFFF) - 1
    m.spare[0, 0] = top + 1
    _problems(m.check(), "free superblock %d is in 2 free lists" % top)


def test_attached_superblock_also_pooled(images):
    m = Img(images["run"])
    c, k, sb = m.owned()[0]
    m.spare[c, int(m.h["spare_slots"]) - 1] = 0
    m.spare[c, 0] = sb + 1
    _problems(m.check(), "superblock %d of class %d is also in a free list" % (sb, k))


def test_stale_shard_count_and_mask(images):
    m = Img(images["trim"])
    s = m.nonempty_shard()
    m.pool[s, 1] += 1
    _problems(m.check(), "its count says")
    m = Img(images["trim"])
    S = m.npool // 8
    m.pmask[s // S] &= ~np.uint64(1 << (s % S))
    _problems(m.check(), "shard %d is not empty but its mask bit is clear" % s)


def test_page_run_is_consistent(images):
    m = Img(images["trim"])
    m.pbits[0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    m.plen[0] = 64
    m.psum[:, 0] |= np.uint64(1)  # word 0 has no free window at any level
    r = m.check()
    assert r["problems"] == 0, r["first"]
    assert r["page_runs"] == 1 and r["pages_used"] == 64


def test_page_run_without_length(images):
    m = Img(images["trim"])
    m.pbits[1] = np.uint64(0xF0)  # pages 68..71 in use, no run starts there
    _problems(m.check(), "page 68 is in use but no run")


def test_wrong_summary_bits(images):
    m = Img(images["trim"])
    m.psum[3, 0] ^= np.uint64(2)
    _problems(m.check(), "page word 1 level 3")
    m = Img(images["trim"])
    m.psum2[0, 0] ^= np.uint64(1)
    _problems(m.check(), "summary word 0 level 0")


def test_not_an_image():
    with pytest.raises(kma_amd.KmaError):
        kma_amd.image_check(np.zeros(4096, dtype=np.uint8))
    with pytest.raises(kma_amd.KmaError):
        kma_amd.image_check(np.zeros(16, dtype=np.uint8))


def test_truncated_image(images):
    with pytest.raises(kma_amd.KmaError):
        kma_amd.image_check(images["trim"][:1000])


def test_corrupt_table_offsets(images):
    """A header whose tables point past the metadata is refused, not walked."""
    for key, value in (("off_bitmap", 1 << 40), ("nsb", 1 << 33), ("off_spare", None)):
        m = Img(images["trim"])
        h = np.frombuffer(m.b[: natimg.HDR_DTYPE.itemsize].tobytes(), dtype=natimg.HDR_DTYPE).copy()
        h[0][key] = value if value is not None else int(h[0]["off_data"]) - 8
        m.b[: natimg.HDR_DTYPE.itemsize] = np.frombuffer(h.tobytes(), dtype=np.uint8)
        with pytest.raises(kma_amd.KmaError):
            kma_amd.image_check(m.b)
