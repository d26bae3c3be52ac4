"""The bench programs' command lines on the CPU (no GPU is touched: -l lists
the runs, and every error path ends before the device is opened).

kma (tb_kma.c) and clTree (tb_clTree.c) take the reference's options
(test/cl.c:228-307, test/tb_kma.c:322-347, test/tb_clTree.c:392-401): -t
x,y,z, -T file of {x,y,z} tuples, -i [step,]iters, a data-set file for
clTree.  The iteration sweep is the reference's loop (tb_kma.c:250-296):
iters alone without a step; with one, 2, then step, 2*step, ... for steps
above 2 and 2, 2 + step, ... for steps of 1 and 2, all up to iters.  The
expected lists below are worked out by hand from that loop."""
import os
import subprocess

import numpy as np
import pytest

import kma_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "kma_amd", "bin")
CFG = os.path.join(ROOT, "kma_amd", "cfg")


def run(prog, *args):
    exe = os.path.join(BIN, prog)
    assert os.path.exists(exe), "bench programs not built (make -C kma_amd/csrc)"
    p = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=60)
    return p.returncode, p.stdout


def planned(out):
    """(threads, iters) of every listed kma run, per kernel."""
    runs, cur = {}, None
    for line in out.splitlines():
        if line.startswith("-- Executing "):
            cur = line.split()[2]
            runs[cur] = []
        elif " threads " in line and cur:
            f = line.split()
            runs[cur].append((int(f[0]), int(f[2])))
    return runs


@pytest.mark.parametrize("spec,iters", [
    ("100", [100]),
    ("10,50", [2, 10, 20, 30, 40, 50]),
    ("2,9", [2, 4, 6, 8]),
    ("1,4", [2, 3, 4]),
    ("3,2", [2]),
    ("5,1", []),
    ("0", [0]),
])
def test_iteration_sweep(spec, iters):
    rc, out = run("kma", "-l", "-t", "16,16,4", "-i", spec)
    assert rc == 0, out
    runs = planned(out)
    assert list(runs) == ["kma_test_malloc", "kma_test_malloc_lowvar", "kma_test_malloc_highvar"]
    for k, v in runs.items():
        assert v == [(1024, i) for i in iters], (k, out)


def test_default_single_work_item_and_one_kernel():
    rc, out = run("kma", "-l", "-k", "kma_test_malloc_lowvar")
    assert rc == 0, out
    assert planned(out) == {"kma_test_malloc_lowvar": [(1, 100)]}


def test_thread_config_file(tmp_path):
    f = tmp_path / "t.cfg"
    # the reference's files put a blank after each comma; both forms read
    f.write_text("3\n{2, 4, 4}\n{8,8,6}\n{ 16, 12, 22}\n")
    rc, out = run("kma", "-l", "-T", f, "-i", "7", "-k", "kma_test_malloc")
    assert rc == 0, out
    assert planned(out) == {"kma_test_malloc": [(32, 7), (384, 7), (4224, 7)]}


def test_shipped_thread_configs():
    expect = {
        "thread_mi355x_small.cfg": [8, 64, 256, 1024, 2048, 4096],
        "thread_mi355x_full.cfg": [64, 256, 1024, 4096, 8192, 16384, 32768, 65536, 131072, 262144, 524288],
        "thread_mi355x_overfull.cfg": [524288, 1 << 20, 1 << 21, 1 << 22],
    }
    for name, threads in expect.items():
        rc, out = run("kma", "-l", "-T", os.path.join(CFG, name), "-k", "kma_test_malloc")
        assert rc == 0, out
        assert [t for t, _ in planned(out)["kma_test_malloc"]] == threads, name


@pytest.mark.parametrize("args,message", [
    (["-x"], "Error: Unknown option: -x"),
    (["-i"], "Error: Unknown option: -i"),
    (["-i", "a,b"], "Error: Unknown option: -i"),
    (["-t"], "Error: -t requires a {x,y,z} tuple"),
    (["-T"], "Error: -T requires a file name"),
    (["-T", "/nonexistent.cfg"], "Error: file /nonexistent.cfg not found"),
    (["-c"], "no CPU device"),
    (["-b", "100"], "Error: Unknown option: -b"),
    (["-C", "1000"], "Error: Unknown option: -C"),
    (["-C", "4"], "Error: Unknown option: -C"),
    (["-k", "kma_test_nothing"], "Error: Unknown option: -k"),
    (["stray"], "Error: Unknown option: stray"),
])
def test_option_errors(args, message):
    rc, out = run("kma", *args)
    assert rc == 255, out
    assert message in out and "Usage: kma [options]" in out, out


def test_malformed_tuple_prints_usage_only():
    rc, out = run("kma", "-t", "1,2")
    assert rc == 255 and "Usage: kma" in out and "Error" not in out, out


@pytest.mark.parametrize("text", ["3\n{1,1,1}\n{2,2,2}\n", "0\n", "x\n{1,1,1}\n", "1\n(1,1,1)\n", "1\n{1 ,1,1}\n"])
def test_malformed_thread_config_file(tmp_path, text):
    f = tmp_path / "bad.cfg"
    f.write_text(text)
    rc, out = run("kma", "-l", "-T", f)
    assert rc == 255 and "Error: malformed thread config file provided" in out, out


def test_clTree_lists_every_form(tmp_path):
    f = tmp_path / "data.txt"
    kma_amd.graph_write(str(f), kma_amd.graph_generate(100, 500, 3))
    rc, out = run("clTree", "-l", "-t", "1000,1,1", f)
    assert rc == 0, out
    lines = out.splitlines()
    assert [l for l in lines if l.startswith("-- Executing")] == [
        "-- Executing clTree_test --", "-- Executing clTree_test_cache --",
        "-- Executing clTree_test_al --", "-- Executing clTree_test_cache --"]
    assert "KMA-1:" in lines and "Poormans heap:" in lines
    assert sum("Threads: 1000" in l and "(4 work-groups of 256)" in l for l in lines) == 4


@pytest.mark.parametrize("content,message", [
    (None, "Error: No data set file specified"),
    ("3\n{1,2}\n{3,4}\n", "Error: Dataset file malformed"),
    ("x\n", "Error: Dataset file malformed"),
])
def test_clTree_data_set_errors(tmp_path, content, message):
    args = ["-l"]
    if content is not None:
        f = tmp_path / "d.txt"
        f.write_text(content)
        args.append(f)
    rc, out = run("clTree", *args)
    assert rc == 255 and message in out, out


def test_clTree_contexts_option():
    rc, out = run("clTree", "-l", "-C", "1000", "-G", "10,10")
    assert rc == 255 and "Error: Unknown option: -C" in out, out
    rc, out = run("clTree", "-l", "-C", "256", "-G", "10,10")
    assert rc == 0, out


def test_clTree_missing_file():
    rc, out = run("clTree", "-l", "/nonexistent/data.txt")
    assert rc == 255 and "Error: specified file cannot be opened (/nonexistent/data.txt)" in out, out


def test_clTree_generated_data_set():
    rc, out = run("clTree", "-l", "-G", "100,1000,7", "-T", os.path.join(CFG, "thread_mi355x_small.cfg"))
    assert rc == 0, out
    assert sum(l.startswith("Threads: ") for l in out.splitlines()) == 4 * 6
