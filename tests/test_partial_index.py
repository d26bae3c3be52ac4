"""Reuse of partly free superblocks (the partial index, include/kma_device.h)
on the CPU.

tests/cpu_sim/sim_partial.cpp runs the device headers on CPU threads: every
lane keeps blocks of one class with random lifetimes, so superblocks
detached from their heads drain part-way; with the live blocks filling
70-90 % of a small heap, allocation keeps succeeding only if refills attach
such superblocks again.  The reference never reattaches a partly free
superblock (kma.cl:369-371: it is reachable again only once empty), so there
is no reference output to match; the checks are the allocator's invariants
(no NULL below the listing threshold's reach, pattern read-back, no double
free, no watchdog expiry, canonical heap after the trim) and the heap
walker on the image taken with live blocks and listed superblocks present,
then with one mistake planted at a time."""
import os
import subprocess

import numpy as np
import pytest

import kma_amd
from tests import natimg
from tests.test_heap_check import Img, _problems

HERE = os.path.dirname(os.path.abspath(__file__))
SIM_DIR = os.path.join(HERE, "cpu_sim")
SIM = os.path.join(SIM_DIR, "_build", "kma_sim_partial")
LISTED = 1 << 41


@pytest.fixture(scope="module")
def sim():
    subprocess.check_call(["make", "-s", "-C", SIM_DIR])
    return SIM


def _run(sim, args, env=None):
    out = subprocess.run([sim] + [str(a) for a in args], capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    f = out.stdout.split()
    return {k: int(v) for k, v in zip(f[0::2], f[1::2])}, out.stdout


@pytest.mark.parametrize("args", [
    # sb_log2 sblocks nwg rounds seed nctx class_size hold_max ring variant
    (14, 256, 4, 60, 1, 8, 256, 21, 24, 0),  # per-lane calls, ~70 % live
    (14, 128, 2, 30, 1, 8, 256, 21, 24, 1),  # work-group staging, ~70 % live
    (14, 256, 4, 60, 2, 8, 256, 27, 32, 0),  # per-lane calls, ~85 % live
])
def test_partial_superblocks_are_reused(sim, args):
    v, raw = _run(sim, args)
    assert v["bad"] == 0 and v["double_frees"] == 0 and v["waits"] == 0, raw
    assert v["canonical_problems"] == 0, raw
    # no NULL at ~70 % live; at ~85 % one XCD's own superblocks can run
    # short for a moment (four work-groups, one per XCD, random lifetimes)
    assert v["nulls"] <= (0 if v["peak_live"] < 0.75 * v["capacity"] else 0.001 * v["mallocs"]), raw
    assert v["peak_live"] > 0.65 * v["capacity"], raw
    # detached superblocks were listed and attached again many times over
    assert v["reattached"] > v["capacity"] // 64 and v["listed"] >= v["reattached"], raw


def test_near_full_heap_stays_consistent(sim):
    """~95 % live: some NULLs are expected (superblocks with fewer free slots
    than the listing threshold are not reachable), but nothing else."""
    v, raw = _run(sim, (14, 256, 4, 60, 3, 8, 256, 31, 40, 0))
    assert v["bad"] == 0 and v["double_frees"] == 0 and v["waits"] == 0, raw
    assert v["canonical_problems"] == 0, raw
    assert v["peak_live"] > 0.9 * v["capacity"], raw
    assert v["nulls"] < 0.1 * v["mallocs"], raw


@pytest.fixture(scope="module")
def image(sim, tmp_path_factory):
    path = str(tmp_path_factory.mktemp("pimg") / "heap")
    v, raw = _run(sim, (14, 256, 4, 60, 2, 8, 256, 27, 32, 0), env=dict(os.environ, KMA_SIM_DUMP=path))
    return np.fromfile(path + ".run", dtype=np.uint8)


class PImg(Img):
    @property
    def pidx(self):
        return self.view("off_pidx", 8 * int(self.h["nclasses"]) * int(self.h["pidx_words"]), "<u8").reshape(
            8, int(self.h["nclasses"]), -1)

    def listed(self):
        """Listed superblocks: (sb, home XCD)."""
        return [(sb, (int(s) >> 38) & 7) for sb, s in enumerate(self.state) if int(s) & LISTED]


def test_walker_accepts_listed_superblocks(image):
    m = PImg(image)
    r = m.check()
    assert r["problems"] == 0, r["first"]
    assert r["sb_listed"] == len(m.listed()) > 0 and r["orphans"] == 0
    assert r["live_blocks"] > 0


def test_listed_bit_without_index_bit(image):
    m = PImg(image)
    sb, x = m.listed()[0]
    m.pidx[x, 0, sb // 64] &= ~np.uint64(1 << (sb % 64))
    _problems(m.check(), "superblock %d: listed bit 1, in 0 partial indexes" % sb)


def test_index_bit_without_listed_bit(image):
    m = PImg(image)
    sb, x = m.listed()[0]
    m.state[sb] &= ~np.uint64(LISTED)
    _problems(m.check(), "superblock %d: listed bit 0, in 1 partial indexes" % sb)


def test_listed_in_another_xcd(image):
    m = PImg(image)
    sb, x = m.listed()[0]
    y = (x + 1) % 8
    m.pidx[x, 0, sb // 64] &= ~np.uint64(1 << (sb % 64))
    m.pidx[y, 0, sb // 64] |= np.uint64(1 << (sb % 64))
    w0 = (m.nsb + 63) // 64
    m.pidx[y, 0, w0 + (sb // 64) // 64] |= np.uint64(1 << ((sb // 64) % 64))
    m.pidx[y, 0, w0 + (w0 + 63) // 64] |= np.uint64(1)
    _problems(m.check(), "superblock %d of XCD %d class 0 is in the partial index of XCD %d" % (sb, x, y))


def test_listed_and_owned(image):
    m = PImg(image)
    sb, x = m.listed()[0]
    m.state[sb] |= np.uint64(1 << 37)
    _problems(m.check(), "superblock %d is listed and owned" % sb)


def test_summary_bit_clear(image):
    m = PImg(image)
    sb, x = m.listed()[0]
    w0 = (m.nsb + 63) // 64
    m.pidx[x, 0, w0 + (sb // 64) // 64] &= ~np.uint64(1 << ((sb // 64) % 64))
    _problems(m.check(), "word %d is not zero but its summary bit is clear" % (sb // 64))


def test_empty_listed_superblock_is_not_an_orphan(image):
    """A listed superblock whose last block was freed while its lister was
    still setting its index bit stays listed, empty, for the next refill of
    its class: consistent, not an orphan."""
    m = PImg(image)
    sb, x = m.listed()[0]
    n = int(m.state[sb]) & 0xFFFFFFFF
    m.bitmap[sb, :] = 0
    m.state[sb] -= np.uint64(n)
    r = m.check()
    assert r["problems"] == 0 and r["orphans"] == 0, r["first"]
