"""CPU simulation of the work-group collective allocator (no GPU).

tests/cpu_sim builds the library's own device headers against a CPU
stand-in for hip_runtime.h (one OS thread per device thread, barriers for
wavefront exchanges) and runs the batched workload body on a native heap in
host memory.  Every run must end with no NULL, no read-back mismatch, no
double or bad free, no watchdog expiry and the metadata back in canonical
form (all bitmaps zero, all superblocks in the pool)."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_sim")
SIM = os.path.join(HERE, "_build", "kma_sim")


@pytest.fixture(scope="module")
def sim():
    subprocess.check_call(["make", "-s", "-C", HERE])
    return SIM


def _run(sim, *args):
    out = subprocess.run([sim] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    head, rest = out.stdout.split("waits")
    f = head.split()
    vals = dict(zip(f[0::2], f[1::2]))
    w, canon = rest.split("canonical_problems")
    vals["canonical_problems"] = canon.strip()
    return vals, [int(x) for x in w.split()], out.stdout


@pytest.mark.parametrize("args", [
    # sb_log2 sblocks nwg nwi rounds mode seed nctx npool [variant]
    # variant bits 0x30: aggregation width 0 wavefront, 0x10/0x20/0x30
    # work-group staging with 8/16/32-lane claim groups
    (14, 1024, 4, 1024, 3, 1, 1, 64, 16, 0x20),     # one context per work-group
    (14, 65536, 16, 4096, 4, 1, 1, 8, 16, 0x20),    # two work-groups per context
    (12, 4096, 4, 1000, 20, 0, 1, 64, 64, 0x20),    # fixed 64 B, partial last work-group
    (12, 4096, 4, 777, 10, 1, 5, 64, 64, 0x20),     # 4 KiB superblocks, mixed classes
    (14, 64, 8, 2048, 3, 1, 9, 8, 16, 0x20),        # tiny heap: pool exhaustion, NULLs allowed
    (14, 4096, 32, 8192, 4, 1, 4, 16, 32, 0x22),    # collective malloc, per-lane kma_free
    (14, 4096, 32, 8192, 4, 1, 4, 16, 32, 0x21),    # per-lane kma_malloc, collective free
    (14, 1024, 32, 8192, 3, 1, 3, 8, 16, 3),        # per-lane malloc and free, 16 wavefronts per context
    (14, 1024, 8, 2048, 3, 1, 6, 16, 16, 0x00),     # wavefront aggregation
    (12, 4096, 4, 777, 10, 1, 7, 64, 64, 0x00),     # wavefront aggregation, partial last wavefront
    (14, 64, 8, 2048, 3, 1, 8, 8, 16, 0x00),        # wavefront aggregation, pool exhaustion
    (14, 1024, 8, 2048, 3, 1, 10, 16, 16, 0x10),    # 8-lane claim groups
    (14, 1024, 8, 2048, 3, 1, 11, 16, 16, 0x30),    # 32-lane claim groups
    (14, 64, 8, 2048, 3, 1, 12, 8, 16, 0x30),       # 32-lane claim groups, pool exhaustion
])
def test_batched_invariants_in_simulation(sim, args):
    vals, waits, raw = _run(sim, *args)
    nwi, rounds = args[3], args[4]
    assert int(vals["bad"]) == 0, raw
    assert int(vals["double_frees"]) == 0 and int(vals["bad_frees"]) == 0, raw
    assert waits == [0, 0, 0, 0], raw
    assert int(vals["canonical_problems"]) == 0, raw
    nulls = int(vals["nulls"])
    assert int(vals["ops"]) == 2 * nwi * rounds - nulls, raw
    if args[1] >= 1024:
        assert nulls == 0, raw


def test_bit_helpers_brute_force(sim):
    """nth_lane, nth_bit and range_bits of the device header against plain
    loops over every bit (tests/cpu_sim/helpers_main.cpp)."""
    out = subprocess.run([os.path.join(HERE, "_build", "kma_helpers")], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "helpers ok" in out.stdout, out.stdout + out.stderr


@pytest.mark.parametrize("args", [
    # log2 sblocks nwg nwi rounds mode
    (12, 127, 32, 8192, 30, 4),  # kma_test_malloc_highvar on kma_create(127)
    (12, 16, 16, 4096, 30, 2),   # kma_test_malloc, 16 superblocks
    (12, 16, 16, 4096, 30, 3),   # kma_test_malloc_lowvar
    (12, 8, 32, 8192, 20, 4),    # kma_test_malloc_highvar, 8 superblocks
])
def test_reference_layout_in_simulation(sim, args):
    """The reference-layout device allocator (kma_ref_device.h) run by the
    reference's bench kernels on CPU threads, built with the host
    AddressSanitizer (any access outside the heap buffer aborts the run):
    no read-back mismatch, and after the last free the superblock free list
    walked by the oracle is back to its initial S - 1 entries."""
    exe = os.path.join(HERE, "_build", "kma_sim_ref_asan")
    out = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    f = out.stdout.split()
    vals = dict(zip(f[0::2], f[1::2]))
    assert int(vals["bad"]) == 0, out.stdout
    assert int(vals["queue"]) == args[1] - 1, out.stdout
    assert int(vals["giveups"]) == 0, out.stdout
    nwi, rounds = args[3], args[4]
    extra = 0 if args[5] == 2 else 2 * rounds  # work-item 0's 4000-B block
    if int(vals["nulls"]) == 0:  # a NULL ends its work-item early
        assert int(vals["ops"]) == 2 * nwi * rounds + extra, out.stdout


def test_reference_layout_bounded_waits(sim):
    """Every bounded wait of the reference-layout allocator on one simulated
    work-item with a 50 ms watchdog (tests/cpu_sim/sim_ref_giveup.cpp): a
    passed deadline makes malloc return NULL at once and is counted; a claim
    on a bitmap planted full gives up, returns its reserved slot to the
    superblock's free count (state word unchanged) and is counted; an
    exhausted heap returns NULL with no give-up and, after every free, the
    oracle walks all S-1 superblocks in the free queue."""
    out = subprocess.run([os.path.join(HERE, "_build", "kma_sim_ref_giveup")], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    for case in ("deadline", "claim", "exhaust"):
        assert "%s ok" % case in out.stdout, out.stdout


@pytest.mark.parametrize("args", [(1, 0), (4, 4), (8, 10), (4, 6, 5), (4, 6, 300)])
def test_linked_queue_in_simulation(sim, args):
    """The linked pointer queue (kma_lqueue.h) on CPU threads, with the host
    AddressSanitizer: every thread enqueues its item, then `rounds` times
    dequeues one and enqueues it again; the reference's validator
    (tb_clQueue.c:126-157) then walks one chain from head to the node tail
    names, holding every item once and the stub once.  With a third argument
    every call names one of that many queues (hashed per thread and round):
    a wavefront's calls then split into the lanes served together and lanes
    served one by one, as a graph's per-node queues see them."""
    out = subprocess.run([os.path.join(HERE, "_build", "kma_sim_lqueue")] + [str(a) for a in args],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    f = out.stdout.split()
    v = {k: int(x) for k, x in zip(f[0::2], f[1::2])}
    assert v["items"] == 256 * args[0] and v["stubs"] == 1 and v["tail_ok"] == 1, out.stdout
    assert v["cycle"] == 0 and v["outside"] == 0 and v["repeats"] == 0 and v["lost"] == 0, out.stdout
