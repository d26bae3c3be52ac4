"""Host-side generator of batched request streams (kma_stream_create): runs
on the CPU, no device needed.  Checks the structure every run depends on
(frees name live blocks of earlier batches, each once; the drain frees the
rest) and the distributions the workload is defined by."""
import numpy as np
import pytest

import kma_amd


def _structure_problems(req, boff, nbatch_gen, drain):
    probs = []
    n = len(req)
    isf = (req >> 31) == 1
    tgt = (req & 0x7FFFFFFF).astype(np.int64)
    batch_of = np.repeat(np.arange(len(boff) - 1), np.diff(boff.astype(np.int64)))
    if len(batch_of) != n:
        return ["batch offsets do not cover the trace"]
    ft = tgt[isf]
    if (ft >= n).any():
        return ["free names a request past the end"]
    if isf[ft].any():
        probs.append("free names a free")
    if (batch_of[ft] >= batch_of[isf]).any():
        probs.append("free names a block of its own or a later batch")
    cnt = np.bincount(ft, minlength=n)
    if (cnt > 1).any():
        probs.append("a block is freed twice")
    if drain and (cnt[~isf] != 1).any():
        probs.append("after the drain some block is not freed exactly once")
    # before the drain, a free only names a block still live then: its
    # malloc precedes it and no earlier free named it (bincount <= 1 above)
    return probs


def test_stream_structure_and_determinism():
    a = kma_amd.Stream(256, 4096, seed=3)
    b = kma_amd.Stream(256, 4096, seed=3)
    c = kma_amd.Stream(256, 4096, seed=4)
    ra, ba = a.trace()
    rb, bb = b.trace()
    rc, _ = c.trace()
    assert np.array_equal(ra, rb) and np.array_equal(ba, bb)
    assert not np.array_equal(ra[:1000], rc[:1000])
    assert _structure_problems(ra, ba, 256, True) == []
    inf = a.info
    assert inf["nbatch"] == 257 and inf["nreq"] == len(ra) == ba[-1]
    assert inf["nmalloc"] + inf["nfree"] == inf["nreq"]
    assert inf["nmalloc"] == inf["nfree"]  # drained
    sizes = ra[(ra >> 31) == 0]
    assert sizes.min() >= 16 and sizes.max() < 65536
    assert inf["small_bytes"] + inf["large_bytes"] > 0


def test_stream_without_drain_leaves_blocks_live():
    s = kma_amd.Stream(64, 1024, seed=7, drain=False)
    req, boff = s.trace()
    assert s.info["nbatch"] == 64
    assert _structure_problems(req, boff, 64, False) == []
    assert s.info["nmalloc"] > s.info["nfree"]


def test_stream_distributions_full_size():
    """The 4,096-batch workload: Zipf(1.2) batch sizes over 1..65,536,
    40 % frees, sizes log-uniform over 16 B .. 64 KiB."""
    s = kma_amd.Stream(4096, 65536, seed=3)
    req, boff = s.trace()
    bs = np.diff(boff.astype(np.int64))[:-1]  # without the drain batch
    assert bs.min() >= 1 and bs.max() <= 65536
    k = np.arange(1, 65537, dtype=np.float64)
    w = k ** -1.2
    mean = float((k * w).sum() / w.sum())
    # mean of 4,096 draws of a heavy-tailed law: loose, but catches a wrong
    # exponent or range (s = 1.1 gives ~2.6x, s = 1.3 ~0.4x this mean)
    assert 0.6 * mean < bs.mean() < 1.6 * mean, (bs.mean(), mean)
    p1 = float(w[0] / w.sum())
    assert abs((bs == 1).mean() - p1) < 0.03, ((bs == 1).mean(), p1)
    gen = req[: boff[-2]]
    isf = (gen >> 31) == 1
    assert abs(isf.mean() - 0.40) < 0.01
    sizes = gen[~isf].astype(np.int64)
    e = np.floor(np.log2(sizes)).astype(np.int64)
    freq = np.bincount(e, minlength=16)[4:16] / len(sizes)
    assert np.allclose(freq, 1 / 12, atol=0.003), freq
    # within an octave the mantissa is uniform: mean offset ~ half an octave
    frac = (sizes - (1 << e)) / (1 << e)
    assert abs(frac.mean() - 0.5) < 0.01
    assert _structure_problems(req, boff, 4096, True) == []


def test_stream_rejects_bad_specs():
    with pytest.raises(kma_amd.KmaError):
        kma_amd.Stream(16, 64, min_size=24)  # not a power of two
    with pytest.raises(kma_amd.KmaError):
        kma_amd.Stream(0, 64)
    with pytest.raises(kma_amd.KmaError):
        kma_amd.Stream(16, 64, free_pct=101)


def test_stream_heap_config_splits_by_demand():
    s = kma_amd.Stream(512, 8192, seed=3)
    inf = s.info
    for sb in (14, 15, 16):
        for table in (kma_amd.TABLE_POW2, kma_amd.TABLE_HALF):
            c = s.heap_config(1 << 30, sb, table, 14)
            d = kma_amd.config_dict(c)
            cl = d["classes"]
            assert cl[0] == 16 and cl[-1] == 16384 and cl == sorted(set(cl))
            assert all(x % 8 == 0 for x in cl)
            if table == kma_amd.TABLE_POW2:
                assert all(x & (x - 1) == 0 for x in cl)
            else:
                assert 24 in cl and 12288 in cl and len(cl) == 21
            small = d["sblocks"] << sb
            large = d["npages"] << 14
            assert small + large <= 1 << 30 and small + large > (1 << 30) - (1 << sb) - (1 << 14)
            want = inf["small_bytes"] / (inf["small_bytes"] + inf["large_bytes"])
            assert abs(small / (1 << 30) - want) < 0.01
