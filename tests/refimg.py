"""Helpers that read reference-layout heap images (64-bit device layout of
kma.h:71-85 / clIndexedQueue.h:38-51) and build traces.  Pure numpy; used by
the tests only."""
import numpy as np


def u32(img, off):
    return int(np.frombuffer(img[off:off + 4].tobytes(), dtype="<u4")[0])


def u64(img, off):
    return int(np.frombuffer(img[off:off + 8].tobytes(), dtype="<u8")[0])


def q_ptr(i, t):
    # the 11-bit tag starts at bit 22 of a 32-bit word, so its top bit is
    # truncated (clIndexedQueue.cl:89)
    return ((i & 0x1FFFE) | (((t + 1) & 0x7FF) << 22)) & 0xFFFFFFFF


def header_bytes(log2):
    return 32 + 8 * (log2 - 1)


def closed_form_init(log2, S, base):
    """Expected image after clheap_init on a zeroed buffer, derived in closed
    form (independently of the oracle's step-by-step enqueue loop):
    sb[0] is the queue's dummy, sb[1..S-2] follow it, sb[S-1] is never
    queued (kma.cl:42,48-51; clIndexedQueue.cl:93-111,119-166)."""
    hdr = header_bytes(log2)
    img = np.zeros((S << log2) + hdr, dtype=np.uint8)

    def w32(off, v):
        img[off:off + 4] = np.frombuffer(np.uint32(v).tobytes(), dtype=np.uint8)

    def w64(off, v):
        img[off:off + 8] = np.frombuffer(np.uint64(v).tobytes(), dtype=np.uint8)

    w64(0, len(img))
    w64(8, base + hdr)
    w64(16, log2)
    w32(24, q_ptr(2, 0))
    # each enqueue advances the tail tag by one: S-2 enqueues after init
    w32(28, q_ptr(2 * (S - 1), S - 2))
    for i in range(S - 1):
        off = hdr + (i << log2)
        if i == S - 2:
            w32(off, q_ptr(0, 0) if i == 0 else 0)
        elif i == 0:
            w32(off, q_ptr(4, 1))
        else:
            w32(off, q_ptr(2 * (i + 2), 0))
    return img


def ref_canonical_free(img, log2, S):
    """Checks of a reference-layout heap with no live block: every class head
    NULL, every bitmap word of every used superblock zero, every superblock
    that was used has free==0 (queued), and the free queue walks to its tail
    with S-1 nodes.  Returns a list of violations."""
    bad = []
    hdr = header_bytes(log2)
    sbsz = 1 << log2
    for c in range(log2 - 1):
        if u64(img, 32 + 8 * c) != 0:
            bad.append("class head %d not NULL" % c)
    for i in range(S):
        off = hdr + i * sbsz
        state = u32(img, off + 4)
        slots = state >> 16
        if slots == 0:
            continue
        if state & 0xFFFF:
            bad.append("sb %d state %#x not queued" % (i, state))
        for w in range((slots + 31) // 32):
            v = u32(img, off + sbsz - 4 * (w + 1))
            if v:
                bad.append("sb %d bitmap word %d = %#x" % (i, w, v))
    # queue walk, in the manner of tb_clIndexedQueue.c:172-210
    base = u64(img, 8)
    stride = u64(img, 16)

    def node(v):
        k = (v >> 1) & 0xFFFFF
        return None if k == 0 else hdr + ((k - 1) << stride)

    head = node(u32(img, 24))
    tail = node(u32(img, 28))
    n, cur = 0, head
    while cur is not None and n <= S + 1:
        nx = u32(img, cur)
        n += 1
        if (nx & 0x1FFFE) == 0:
            break
        cur = node(nx)
    if cur != tail:
        bad.append("queue end != tail")
    if n != S - 1:
        bad.append("queue holds %d nodes, expected %d" % (n, S - 1))
    del base
    return bad


def random_trace(rng, n, sizes, free_frac=0.45, max_live=None):
    """Op list: >=0 malloc(size), <0 free(op index).  Frees pick a random
    live earlier malloc."""
    ops = []
    live = []
    for _ in range(n):
        if live and (rng.random() < free_frac or (max_live and len(live) >= max_live)):
            j = live.pop(rng.integers(len(live)))
            ops.append(-j - 1)
        else:
            live.append(len(ops))
            ops.append(int(rng.choice(sizes)))
    return np.array(ops, dtype=np.int64), live


def free_all(ops, live):
    tail = [-j - 1 for j in live]
    return np.concatenate([ops, np.array(tail, dtype=np.int64)])
