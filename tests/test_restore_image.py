"""Byte comparison of a native heap's metadata with its initial image.

The allocator's invariant: after a trace that frees every block, followed by
kma_trim, the metadata image equals that of a freshly created heap in every
bit outside a documented exclusion mask (tests/natimg.py: statistics, hint
and request words, superblock generations, pool heads and links, stale page
lengths).  tests/cpu_sim/sim_partial.cpp writes the fresh image, the image
with live blocks and the image after the final frees and the trim; the
comparison must find nothing in the last, and it must find every byte
planted outside the mask (and none inside it)."""
import os
import subprocess

import numpy as np
import pytest

from tests import natimg

HERE = os.path.dirname(os.path.abspath(__file__))
SIM_DIR = os.path.join(HERE, "cpu_sim")
SIM = os.path.join(SIM_DIR, "_build", "kma_sim_partial")


@pytest.fixture(scope="module")
def images(tmp_path_factory):
    subprocess.check_call(["make", "-s", "-C", SIM_DIR])
    path = str(tmp_path_factory.mktemp("rimg") / "heap")
    out = subprocess.run([SIM, "14", "256", "4", "60", "2", "8", "256", "27", "32", "0"], capture_output=True,
                         text=True, timeout=300, env=dict(os.environ, KMA_SIM_DUMP=path))
    assert out.returncode == 0, out.stdout + out.stderr
    load = lambda s: np.fromfile(path + s, dtype=np.uint8)
    return load(".init"), load(".run"), load(".final")


def _hdr(img):
    return np.frombuffer(img[: natimg.HDR_DTYPE.itemsize].tobytes(), dtype=natimg.HDR_DTYPE)[0]


def test_restored_after_full_trace(images):
    init, run, final = images
    assert natimg.restored_diff(init, final) == []
    # the comparison does see a heap in use: live blocks' bitmaps and counts
    tables = {t for _, t, _, _ in natimg.restored_diff(init, run)}
    assert "state" in tables


def test_mask_covers_only_the_documented_fields(images):
    """Bits inside the mask may differ; a byte changed anywhere else in each
    table is reported, with its table's name."""
    init, _, final = images
    h = _hdr(init)
    nsb, npool = int(h["nsb"]), int(h["npool"])
    masked = final.copy()
    so, st = int(h["off_stats"]), int(h["off_state"])
    masked[so + 17] ^= 0xFF                          # a statistics counter
    masked[so + 256 + 128 * 3 + 8] ^= 0xFF           # XCD 3's clean-request word
    masked[st + 8 * 5 + 7] ^= 0x80                    # superblock 5's generation
    masked[st + 8 * 5 + 5] ^= 0x04                    # bit 42: generation
    masked[int(h["off_next"]) + 4 * 7] ^= 0x01       # a pool link
    masked[int(h["off_pool"]) + 128 * (npool - 1)] ^= 0x01  # a shard head
    assert natimg.restored_diff(init, masked) == []
    cases = {
        "state": st + 8 * 9,                          # superblock 9's count
        "bitmap": int(h["off_bitmap"]) + 4 * int(h["bm_words"]) * 3,
        "heads": int(h["off_heads"]) + 8 * 40,
        "pool": int(h["off_pool"]) + 128 * 2 + 16,    # past a shard's head and count
        "spare": int(h["off_spare"]) + 4,
        "pidx": int(h["off_pidx"]) + 8,
        "stats": so + 256 + 128 * 2 + 40,             # XCD 2's line past its four words
        "header": 24,
    }
    for table, off in cases.items():
        img = final.copy()
        img[off] ^= 0x01
        d = natimg.restored_diff(init, img)
        assert [(o, t) for o, t, _, _ in d] == [(off, table)], (table, d)
    # bit 41 (listed) and bit 40 (home XCD) are outside the mask
    for bit in (0x01, 0x02):
        img = final.copy()
        img[st + 8 * 11 + 5] ^= bit
        assert [(o, t) for o, t, _, _ in natimg.restored_diff(init, img)] == [(st + 8 * 11 + 5, "state")]
    assert nsb > 11
