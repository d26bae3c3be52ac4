"""The bench programs on the GPU: `kma` runs the reference's three test
kernels over thread configurations and iteration counts, `clTree` builds a
graph in the reference's four forms; every run line is printed and the
device counters (-g) agree with what the kernels do: kma_test_malloc issues
one malloc and one free per work-item per iteration (kma.cl:424-461); the
_lowvar and _highvar kernels add a 4000-B malloc and free on work-item 0
(kma.cl:463-526)."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "kma_amd", "bin")
CFG = os.path.join(ROOT, "kma_amd", "cfg")
LINE = re.compile(r"^(\d+)\s+threads (\d+)\s+iters: Avg: (\d+\.\d{6}) \( (\d+\.\d{6}) - (\d+\.\d{6}) \) Total: (\d+\.\d{6})$")


def run(prog, *args, timeout=100):
    print("bench program: %s %s" % (prog, " ".join(map(str, args))), flush=True)
    p = subprocess.run([os.path.join(BIN, prog)] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=timeout)
    return p.returncode, p.stdout


def parse(out):
    """{kernel: [(threads, iters, avg_s, counters dict)]}"""
    res, cur = {}, None
    for line in out.splitlines():
        if line.startswith("-- Executing "):
            cur = line.split()[2]
            res[cur] = []
            continue
        m = LINE.match(line)
        if m:
            res[cur].append([int(m.group(1)), int(m.group(2)), float(m.group(3)), {}])
        elif line.startswith("\tops "):
            f = line.split()
            w = f.index("waits")
            c = {f[i]: float(f[i + 1]) for i in range(0, w, 2)}
            c["waits"] = [int(x) for x in f[w + 1: w + 5]]
            c[f[w + 5]] = float(f[w + 6])
            res[cur][-1][3] = c
    return res


def _ops(kernel, threads, iters):
    return 2 * threads * iters + (0 if kernel == "kma_test_malloc" else 2 * iters)


def test_kma_reference_layout_sweep():
    rc, out = run("kma", "-g", "-t", "16,16,4", "-i", "10,30")
    assert rc == 0, out
    res = parse(out)
    assert list(res) == ["kma_test_malloc", "kma_test_malloc_lowvar", "kma_test_malloc_highvar"], out
    for k, rows in res.items():
        assert [(t, i) for t, i, _, _ in rows] == [(1024, 2), (1024, 10), (1024, 20), (1024, 30)], out
        for t, i, avg, c in rows:
            assert c["nulls"] == 0 and c["bad"] == 0 and c["waits"] == [0, 0, 0, 0], (k, c)
            assert c["ops"] == _ops(k, t, i), (k, t, i, c)
            assert 0 < avg < 1.0


def test_kma_native_full_sweep():
    """The MI355X sweep up to full occupancy (524,288 work-items) on a native
    heap, with the heap walker after every run."""
    rc, out = run("kma", "-N", "-g", "-T", os.path.join(CFG, "thread_mi355x_full.cfg"), "-i", "20",
                  "-k", "kma_test_malloc_highvar")
    assert rc == 0, out
    rows = parse(out)["kma_test_malloc_highvar"]
    assert [t for t, _, _, _ in rows][-1] == 524288 and len(rows) == 11
    for t, i, _, c in rows:
        assert c["nulls"] == 0 and c["bad"] == 0 and c["double_frees"] == 0, c
        assert c["ops"] == _ops("kma_test_malloc_highvar", t, i), (t, c)
    walks = [l for l in out.splitlines() if l.startswith("\theap walk:")]
    assert len(walks) == 11 and all(" 0 problems, 0 orphans, 0 live blocks" in w for w in walks), walks


def test_kma_reference_layout_overfull():
    """32,768 work-items of the high-variance kernel on the 127-superblock
    heap, which cannot hold one block of every work-item at once (about
    200 superblocks' worth): work-items that get NULL stop (kma.cl:507-509);
    the run completes, prints, and no block is handed out twice (bad = 0)."""
    rc, out = run("kma", "-g", "-t", "32,32,32", "-i", "20", "-k", "kma_test_malloc_highvar")
    assert rc == 0, out
    rows = parse(out)["kma_test_malloc_highvar"]
    assert len(rows) == 1 and rows[0][0] == 32768
    c = rows[0][3]
    assert c["bad"] == 0 and c["mallocs"] >= c["nulls"], c


def test_clTree_all_forms():
    rc, out = run("clTree", "-g", "-G", "3000,20000,5", "-T", os.path.join(CFG, "thread_mi355x_small.cfg"))
    assert rc == 0, out
    lines = out.splitlines()
    runs = [l for l in lines if l.startswith("Threads: ")]
    assert len(runs) == 4 * 6, out
    stats = [l for l in lines if l.startswith("\tedges ")]
    assert len(stats) == 4 * 6
    for s in stats:
        f = s.split()
        d = {f[i]: f[i + 1] for i in range(0, len(f) - 1, 2)}
        assert d["edges"] == "20000" and d["links_found"] == "20000" and d["nulls"] == "0", s


def test_kma_print_buffer():
    """-p dumps the first 16 superblocks of the heap as 32-bit words, four to
    a line with the byte offset (tb_kma.c:296-306).  On kma_create(127) the
    first word is the heap's size in bytes: 127 superblocks of 4 KiB plus a
    header of 32 + 8 * 11 bytes (kma.c:79-127) = 520,312 = 0x7f078."""
    rc, out = run("kma", "-p", "-t", "64,1,1", "-i", "2", "-k", "kma_test_malloc")
    assert rc == 0, out
    assert "00000000: 0007f078 00000000 " in out, out[:2000]
    lines = [l for l in out.splitlines() if re.match(r"^[0-9a-f]{8}: ", l)]
    assert len(lines) == 16 * 4096 // 16, len(lines)
