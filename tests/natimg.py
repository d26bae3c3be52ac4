"""Decoder for native-layout heap images (kma_dheap in include/kma_device.h).
Used by the tests only."""
import numpy as np

HDR_DTYPE = np.dtype([
    ("magic", "<u8"), ("bytes", "<u8"), ("nsb", "<u8"), ("npages", "<u8"),
    ("sb_log2", "<u4"), ("page_log2", "<u4"), ("nclasses", "<u4"), ("nctx", "<u4"),
    ("npool", "<u4"), ("bm_words", "<u4"), ("max_small", "<u4"), ("pad0", "<u4"),
    ("off_state", "<u8"), ("off_next", "<u8"), ("off_bitmap", "<u8"), ("off_heads", "<u8"),
    ("off_pool", "<u8"), ("off_pbitmap", "<u8"), ("off_plen", "<u8"), ("off_data", "<u8"),
    ("off_pages", "<u8"), ("off_stats", "<u8"),
    ("class_size", "<u4", 32), ("class_cap", "<u4", 32), ("class_magic", "<u8", 32),
    ("class_for_log2", "u1", 64),
    ("off_psum", "<u8"), ("off_spare", "<u8"), ("spare_slots", "<u4"), ("pad1", "<u4"),
    ("off_pidx", "<u8"), ("pidx_words", "<u4"), ("no_partial", "<u4"),
])

MAGIC = 0x4E414D4B31414D4B
PAGE_LEVELS = 7
HEADS_PER_CTX = 32
POOL_STRIDE = 16  # u64 words


class NatHeap:
    def __init__(self, heap):
        raw = heap.snapshot(0, HDR_DTYPE.itemsize)
        self.h = np.frombuffer(raw.tobytes(), dtype=HDR_DTYPE)[0]
        assert int(self.h["magic"]) == MAGIC
        self.heap = heap
        self.nsb = int(self.h["nsb"])
        self.npages = int(self.h["npages"])
        self.sb_log2 = int(self.h["sb_log2"])
        self.page_log2 = int(self.h["page_log2"])
        self.nclasses = int(self.h["nclasses"])
        self.nctx = int(self.h["nctx"])
        self.npool = int(self.h["npool"])
        self.bm_words = int(self.h["bm_words"])
        self.max_small = int(self.h["max_small"])
        self.classes = [int(x) for x in self.h["class_size"][: self.nclasses]]
        self.caps = [int(x) for x in self.h["class_cap"][: self.nclasses]]
        self.off_data = int(self.h["off_data"])
        self.off_pages = int(self.h["off_pages"])

    def meta(self):
        """Metadata snapshot (everything before the superblock data)."""
        img = self.heap.snapshot(0, int(self.h["off_data"]))
        h = self.h

        def arr(off, n, dt):
            off = int(off)
            return np.frombuffer(img[off: off + n * np.dtype(dt).itemsize].tobytes(), dtype=dt)

        nwords = (self.npages + 63) // 64
        ns = (nwords + 63) // 64
        return {
            "stats": arr(h["off_stats"], 8, "<u8"),
            "pmask": arr(int(h["off_stats"]) + 256, 8 * 16, "<u8")[::16],
            "psum": arr(h["off_psum"], PAGE_LEVELS * ns, "<u8").reshape(PAGE_LEVELS, ns),
            "psum2": arr(int(h["off_psum"]) + PAGE_LEVELS * ns * 8, PAGE_LEVELS * ((ns + 63) // 64),
                         "<u8").reshape(PAGE_LEVELS, (ns + 63) // 64),
            "spare": arr(h["off_spare"], self.nctx * 32, "<u4").reshape(self.nctx, 32),
            "state": arr(h["off_state"], self.nsb, "<u8"),
            "next": arr(h["off_next"], self.nsb, "<u4"),
            "bitmap": arr(h["off_bitmap"], self.nsb * self.bm_words, "<u4"),
            "heads": arr(h["off_heads"], self.nctx * HEADS_PER_CTX, "<u8"),
            "pool": arr(h["off_pool"], self.npool * POOL_STRIDE, "<u8").reshape(self.npool, POOL_STRIDE),
            "pbitmap": arr(h["off_pbitmap"], (self.npages + 63) // 64, "<u8"),
            "plen": arr(h["off_plen"], self.npages, "u1"),
            "pidx": arr(h["off_pidx"], 8 * self.nclasses * int(h["pidx_words"]), "<u8"),
        }

    def pool_members(self, m):
        """Walk every pool shard; returns (list of superblocks, problems)."""
        seen, probs = [], []
        for s in range(self.npool):
            top = int(m["pool"][s, 0]) & 0xFFFFFFFF
            n = 0
            while top:
                seen.append(top - 1)
                top = int(m["next"][top - 1])
                n += 1
                if n > self.nsb:
                    probs.append("cycle in shard %d" % s)
                    break
            if n != int(m["pool"][s, 1]):
                probs.append("shard %d count %d != walked %d" % (s, int(m["pool"][s, 1]), n))
        return seen, probs

    def canonical_problems(self):
        """A heap with no live block, after trim: every bitmap and page-bitmap
        word zero, every state word zero apart from its generation, every
        context head, spare line and partial index empty,
        every superblock in exactly one pool shard."""
        m = self.meta()
        probs = []
        if m["spare"].any():
            probs.append("%d context spare slots hold superblocks" % int(np.count_nonzero(m["spare"])))
        if m["bitmap"].any():
            probs.append("%d bitmap words nonzero" % int(np.count_nonzero(m["bitmap"])))
        if m["pbitmap"].any():
            probs.append("page bitmap nonzero")
        gen_mask = np.uint64((1 << 42) - 1)  # all but the generation
        if (m["state"] & gen_mask).any():
            probs.append("%d state words not free" % int(np.count_nonzero(m["state"] & gen_mask)))
        if m["heads"].any():
            probs.append("context heads not empty")
        if m["pidx"].any():
            probs.append("%d partial-index words not zero" % int(np.count_nonzero(m["pidx"])))
        seen, p = self.pool_members(m)
        probs += p
        if sorted(seen) != list(range(self.nsb)):
            probs.append("pool holds %d superblocks (%d distinct), expected %d"
                         % (len(seen), len(set(seen)), self.nsb))
        per = self.npool // 8
        for sh in range(self.npool):
            if int(m["pool"][sh, 0]) & 0xFFFFFFFF and not (int(m["pmask"][sh // per]) >> (sh % per)) & 1:
                probs.append("shard %d non-empty but its hint bit is clear" % sh)
        if self.npages and not np.array_equal(m["psum"], self.psum_initial()):
            probs.append("page-window summaries differ from the empty-heap image")
        if self.npages and not np.array_equal(m["psum2"], self.psum2_initial()):
            probs.append("top-level page-window summaries differ from the empty-heap image")
        return probs

    def psum2_initial(self):
        """Top-level summary bits of an empty page region: set for summary
        words that do not exist or are all ones."""
        ps = self.psum_initial()
        ns = ps.shape[1]
        ns2 = (ns + 63) // 64
        out = np.zeros((PAGE_LEVELS, ns2), dtype=np.uint64)
        for lv in range(PAGE_LEVELS):
            for si in range(ns2 * 64):
                if si >= ns or int(ps[lv, si]) == (1 << 64) - 1:
                    out[lv, si // 64] |= np.uint64(1 << (si % 64))
        return out

    def psum_initial(self):
        """Page-window summary bits of an empty page region: set only for
        words that do not exist or cannot hold a window of the level."""
        nwords = (self.npages + 63) // 64
        ns = (nwords + 63) // 64
        out = np.zeros((PAGE_LEVELS, ns), dtype=np.uint64)
        for lv in range(PAGE_LEVELS):
            for wi in range(ns * 64):
                valid = min(64, max(0, self.npages - 64 * wi))
                if valid < (1 << lv):  # no aligned window of 2^lv valid pages
                    out[lv, wi // 64] |= np.uint64(1 << (wi % 64))
        return out


def check_blocks(nh, offs, sizes):
    """Invariants of live native blocks: in-heap, size-class-correct, aligned
    to the block size, non-overlapping.  offs are heap offsets."""
    probs = []
    spans = []
    data_end = nh.off_data + (nh.nsb << nh.sb_log2)
    for off, size in zip(offs, sizes):
        off, size = int(off), int(size)
        if off == 0:
            continue
        if size <= nh.max_small:
            cls = next(c for c in nh.classes if c >= size)
            if not (nh.off_data <= off < data_end):
                probs.append("block %#x outside superblock data" % off)
                continue
            inner = (off - nh.off_data) & ((1 << nh.sb_log2) - 1)
            if inner % cls:
                probs.append("block %#x not aligned to class %d" % (off, cls))
            if inner + cls > (1 << nh.sb_log2):
                probs.append("block %#x crosses its superblock" % off)
            spans.append((off, off + cls))
        else:
            psz = 1 << nh.page_log2
            if not (nh.off_pages <= off < nh.off_pages + (nh.npages << nh.page_log2)):
                probs.append("large block %#x outside page region" % off)
                continue
            if (off - nh.off_pages) % psz:
                probs.append("large block %#x not page aligned" % off)
            spans.append((off, off + ((size + psz - 1) // psz) * psz))
    spans.sort()
    for a, b in zip(spans, spans[1:]):
        if a[1] > b[0]:
            probs.append("overlap %#x..%#x / %#x" % (a[0], a[1], b[0]))
    return probs


# ---- byte comparison with the initial image ----
#
# After a trace that frees every block, followed by kma_trim, the metadata
# image must equal the image of a freshly created heap in every bit except
# these (the exclusion mask), each of which legitimately carries history:
#   * the statistics counters (256 B at off_stats);
#   * per XCD (128-B lines after them): the shard hint mask, the clean
#     request word, the exhaustion window and the pool-pressure word (words
#     0..3; the rest of each line must match);
#   * the generation of every superblock state word (bits 42..63: the ABA
#     guard advances on every release);
#   * the pool: each shard's {tag, top} head and count (words 0 and 1 of its
#     128-B line; the rest must match) and the pool links (which shard holds
#     which superblock, in which order, follows the release order);
#   * the page-run lengths (a freed run's length stays behind; it is read
#     only where its run's first bit is set).
# The excluded pool structure is checked separately: every superblock in
# exactly one shard, every non-empty shard with its hint bit
# (NatHeap.canonical_problems).

def restore_mask(h):
    """Per-byte bit mask over the metadata image [0, off_data) of header `h`
    (HDR_DTYPE): set bits may differ from the initial image."""
    n = int(h["off_data"])
    m = np.zeros(n, np.uint8)
    so = int(h["off_stats"])
    m[so: so + 256] = 0xFF
    for x in range(8):
        line = so + 256 + 128 * x
        m[line: line + 32] = 0xFF
    nsb = int(h["nsb"])
    st = m[int(h["off_state"]): int(h["off_state"]) + 8 * nsb].reshape(nsb, 8)
    st[:, 5] |= 0xFC  # bits 42..47
    st[:, 6:8] = 0xFF  # bits 48..63
    m[int(h["off_next"]): int(h["off_next"]) + 4 * nsb] = 0xFF
    npool = int(h["npool"])
    pool = m[int(h["off_pool"]): int(h["off_pool"]) + 128 * npool].reshape(npool, 128)
    pool[:, 0:16] = 0xFF
    m[int(h["off_plen"]): int(h["off_plen"]) + int(h["npages"])] = 0xFF
    return m


def _region(h, off):
    """Name of the table that holds byte `off` of the metadata image."""
    keys = ["off_stats", "off_state", "off_next", "off_bitmap", "off_heads", "off_pool", "off_pbitmap",
            "off_plen", "off_psum", "off_spare", "off_pidx", "off_data"]
    starts = sorted((int(h[k]), k[4:]) for k in keys)
    name = "header"
    for s, k in starts:
        if off >= s:
            name = k
    return name


def restored_diff(img0, img1):
    """Bytes of metadata image img1 that differ from the initial image img0
    outside the exclusion mask: a list of (offset, table, initial, now), at
    most 20.  Both are uint8 arrays of at least off_data bytes."""
    h = np.frombuffer(img0[: HDR_DTYPE.itemsize].tobytes(), dtype=HDR_DTYPE)[0]
    n = int(h["off_data"])
    a, b = np.asarray(img0[:n], np.uint8), np.asarray(img1[:n], np.uint8)
    diff = (a ^ b) & ~restore_mask(h)
    idx = np.nonzero(diff)[0]
    return [(int(i), _region(h, int(i)), int(a[i]), int(b[i])) for i in idx[:20]]
