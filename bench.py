#!/usr/bin/env python3
"""Benchmark: heap ops/s (mallocs + frees per second) of the MI355X device
allocator at 524,288 concurrent work-items, batched persistent form vs the
plain per-launch form.

Workload (full occupancy): 1 GiB native heap of 16 KiB superblocks;
524,288 work-items (2,048 work-groups x 256); each round every work-item
mallocs 2^e bytes (e uniform in 3..11, splitmix64 of seed 1), holds it for
0..64 iterations of one u32 write, reads its pattern back and frees it.
One step = one round = 1,048,576 heap ops.  Inputs are generated on the
device from the seed; nothing crosses PCIe in the timed region.

Also measured (key "batched_stream"): the batched-stream workload, a
host-generated trace of 4,096 uneven batches already resident in HBM, run
as one launch per batch (plain) and as one graph-captured persistent launch
with the autotuned heap.  Key "serving": the device-resident serving loop
(10 s of Poisson arrivals on a 64 GiB heap) with malloc latency
percentiles.

Prints ONE JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "mallocs+frees/s at 512K work-items, batched vs plain; p99 alloc latency"
NWI = 524288
BLOCK = 256
SB_LOG2 = 14
SBLOCKS = 65536  # 1 GiB
# Allocation contexts of the full-occupancy heap: 4,096 = one per resident
# wavefront (256 CUs x 4 SIMDs x 4 waves at the kernels' register use), so
# per-lane calls of different wavefronts do not share a context's heads.
# The library's default (a quarter of the heap in heads at most) gives 1,024.
NCTX = 4096
SEED = 1
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30  # the copy probe: reads and writes 1 GiB
# Counter-measured L2 requests and HBM traffic of the batched launch, per
# round (FETCH_SIZE / WRITE_SIZE scaled by the copy probe's reading).
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "counters_full_occupancy_batched.md")
L2_PROFILE = TRAFFIC_PROFILE


def l2_requests_per_round():
    """L2 (TCC) requests per round from the committed counter profile, or None."""
    try:
        for line in open(L2_PROFILE):
            if line.startswith("tcc_requests_per_round:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def l2_roofline(ms_per_round, store_peak):
    """L2 request rate of the batched launch (counter-measured requests per
    round over the measured time per round) against the scattered-store
    rate of the same shape."""
    req = l2_requests_per_round()
    if not req:
        return None
    rate = req / (ms_per_round * 1e-3)
    return {"requests_per_round": req, "achieved_requests_per_s": rate,
            "peak_scattered_4b_stores_per_s": store_peak, "frac": rate / store_peak,
            "source": "profiles/counters_full_occupancy_batched.md"}


def traffic_per_round():
    """Bytes per round from the committed counter profile, or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_round:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None
# Algorithmic bytes per heap op: one 4-B bitmap word read-modify-write per
# malloc or free (8 B), plus the workload payload per work-item-round
# (pattern write + mean 32 hold writes + read-back, 4 B each) split over
# its two ops: (1 + 32 + 1) * 4 / 2 = 68 B.
ALGO_BYTES_PER_OP = 8 + 68


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(rounds=4, runs=5):
    """Reference algorithm, CPU restatement (oracle/refmodel.c, pthreads with
    GCC atomics, same CAS protocol), one thread per host core (nproc), each
    thread running a contiguous slice of the work-items as a CPU OpenCL
    runtime runs a barrier-free kernel; the median of `runs` runs."""
    from oracle import refmodel as rm

    threads = max(1, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1))
    res = [rm.cpu_run(SB_LOG2, SBLOCKS, threads, NWI, rounds, 1, 0, SEED) for _ in range(runs)]
    rates = sorted(r["ops"] / r["seconds"] for r in res)
    med = rates[len(rates) // 2]
    return {
        "value": med,
        "unit": "heap ops/s",
        "cores": threads,
        "kind": "port",
        "sample": "median of %d runs (%s heap ops/s) of rounds 0-%d of the full-occupancy trace, all %d "
                  "work-items (%d ops per run, %.1f s in all) on %d host threads (nproc); reference protocol "
                  "incl. its kma.cl:338 write-after-release (%d read-back mismatches in all)"
                  % (runs, ", ".join("%.3g" % x for x in rates), rounds - 1, NWI, res[0]["ops"],
                     sum(r["seconds"] for r in res), threads, sum(r["bad"] for r in res)),
    }


# Batched-stream workload: 4,096 batches, Zipf(1.2) sizes over 1..65,536
# requests, 40 % frees of live blocks of earlier batches, sizes log-uniform
# 16 B .. 64 KiB (>= 16 KiB on pages), 4 GiB heap, seed 3, drained at the end.
STREAM = dict(nbatch=4096, max_batch=65536, zipf=1.2, free_pct=40, min_size=16, max_size=65536,
              page_min=16384, seed=3)
STREAM_HEAP = 4 << 30


def _stream_line(st):
    good = st["mallocs"] + st["frees"]
    return {
        "ops_per_s": st["ops"] / (st["gpu_ms"] * 1e-3),
        "successful_ops_per_s": good / (st["gpu_ms"] * 1e-3),
        "ms_per_replay": st["gpu_ms"],
        "launches": st["launches"],
        "ops": st["ops"],
        "nulls": st["nulls"],
        "large_nulls": st["large_nulls"],
        "violations": int(st["bad"] + st["wait_timeouts"] + st["double_frees"] + st["bad_frees"] + sum(st["waits"])),
    }


def bench_stream(reps):
    """Per-launch form (one launch per batch) on the default heap, then the
    autotuned persistent form (one graph launch per replay)."""
    import kma_amd

    s = kma_amd.Stream(**STREAM)
    log("bench: stream of %d requests in %d batches" % (s.info["nreq"], s.info["nbatch"]))
    cfg = s.heap_config(STREAM_HEAP, 14, kma_amd.TABLE_POW2, 14)
    with kma_amd.Heap.from_config(cfg) as h:
        plain = _stream_line(s.run(h, form=kma_amd.FORM_PLAIN, warmup=1, reps=max(1, reps // 2)))
        log("bench: stream plain %s" % plain)
        untuned = _stream_line(s.run(h, form=kma_amd.FORM_BATCHED, nwg=1024, per_thread=1, warmup=1, reps=reps))
        log("bench: stream persistent, default heap %s" % untuned)
    # the persistent form may not buy speed with NULLs: no more than the
    # per-launch form returned
    limit = plain["nulls"]
    tune = s.autotune(STREAM_HEAP, 14, null_limit=limit, reps=2)
    log("bench: autotuner picked sb_log2 %d table %d agg %#x nwg %d per_thread %d (%.3e ops/s, nulls %d) of %d trials"
        % (tune["config"]["sb_log2"], tune["table"], tune["agg"], tune["nwg"], tune["per_thread"], tune["ops_per_s"],
           tune["nulls"], len(tune["trials"])))
    with kma_amd.Heap.from_config(tune["cfg"]) as h:
        batched = _stream_line(s.run(h, form=kma_amd.FORM_BATCHED, nwg=tune["nwg"],
                                     per_thread=tune["per_thread"], warmup=1, reps=reps, variant=tune["agg"]))
    log("bench: stream persistent, tuned %s" % batched)
    return {
        "workload": "4,096 batches, Zipf(1.2) over 1..65,536 requests, 40% frees of earlier batches' "
                    "live blocks, sizes log-uniform 16 B..64 KiB (>16 KiB on 16 KiB pages), 4 GiB heap, "
                    "seed 3, drain batch at the end",
        "requests": s.info["nreq"],
        "batches": s.info["nbatch"],
        "peak_live_demand_bytes": s.info["peak_small_bytes"] + s.info["peak_large_bytes"],
        "plain": plain,
        "persistent_default_heap": untuned,
        "persistent_tuned": batched,
        "batched_over_plain": batched["ops_per_s"] / plain["ops_per_s"],
        "null_limit": limit,
        "tuned": {"sb_log2": tune["config"]["sb_log2"], "classes": tune["config"]["classes"],
                  "sblocks": tune["config"]["sblocks"], "npages": tune["config"]["npages"],
                  "agg": tune["agg"], "nwg": tune["nwg"], "per_thread": tune["per_thread"],
                  "no_partial": tune["config"]["no_partial"]},
        "trials": [{k: t[k] for k in ("sb_log2", "table", "agg", "nwg", "per_thread", "no_partial", "ops_per_s", "nulls",
                                      "violations")}
                   for t in tune["trials"]],
    }


# Serving loop: Poisson arrivals (2,000 jobs/ms), log-normal lengths
# (mu 6, sigma 1 steps, capped at 8,192), one 16 KiB block per live job per
# step, whole chain freed at completion, 64 GiB heap (1 MiB superblocks of
# 64 blocks), admission within 90 % of the heap, 10 s, seed 4.
SERVE = dict(rate_per_ms=2000.0, len_mu=6.0, len_sigma=1.0, duration_ms=10000.0, max_len=8192,
             block_bytes=16384, watermark_pct=90, max_live=32768, seed=4)
SERVE_HEAP_SB = 65536  # x 1 MiB = 64 GiB


GRAPH = dict(nnodes=200000, nedges=2000000, seed=21)


def bench_graph():
    """Graph building (the reference's tree/queue application) on a 1 GiB
    native heap: every node and link form, plus the cache form on the bump
    allocator; edges linked per second of the build launch."""
    import kma_amd

    edges = kma_amd.graph_generate(GRAPH["nnodes"], GRAPH["nedges"], GRAPH["seed"])
    out = {"workload": "%d edges over %d node ids (seed %d), 1 GiB heap of 16 KiB superblocks with %d "
                       "allocation contexts, 1,024 work-groups of 256"
                       % (GRAPH["nedges"], GRAPH["nnodes"], GRAPH["seed"], NCTX)}
    for mode, name in ((kma_amd.GRAPH_PLAIN, "plain"), (kma_amd.GRAPH_CACHE, "cache"),
                       (kma_amd.GRAPH_COLLECTIVE, "collective")):
        with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, SB_LOG2, SBLOCKS, nctx=NCTX) as h:
            st = kma_amd.graph_build(edges, mode, heap=h, nwg=1024)
        out[name] = {"edges_per_s": st["edges"] / (st["gpu_ms"] * 1e-3), "ms": st["gpu_ms"],
                     "violations": int(st["violations"] + st["bad"] + st["nulls"])}
        log("bench: graph %s %s" % (name, out[name]))
    with kma_amd.Pma(1 << 30, 16) as p:
        st = kma_amd.graph_build(edges, kma_amd.GRAPH_CACHE, pma=p, nwg=1024)
    out["cache_bump_allocator"] = {"edges_per_s": st["edges"] / (st["gpu_ms"] * 1e-3), "ms": st["gpu_ms"],
                                   "violations": int(st["violations"] + st["bad"] + st["nulls"])}
    log("bench: graph cache on the bump allocator %s" % out["cache_bump_allocator"])
    return out


def bench_serve(duration_ms, admit):
    import kma_amd

    spec = kma_amd.serve_spec(**dict(SERVE, duration_ms=duration_ms, admit=admit))
    with kma_amd.Heap(layout=kma_amd.LAYOUT_NATIVE, sb_log2=20, sblocks=SERVE_HEAP_SB, classes=[16384]) as h:
        st, _ = kma_amd.serve_run(h, spec)
    policy = "projected" if admit == kma_amd.ADMIT_PROJECTED else "watermark"
    log("bench: serving loop, %s admission %s" % (policy, {k: (round(v, 3) if isinstance(v, float) else v)
                                                            for k, v in st.items()}))
    return {
        "workload": "Poisson arrivals 2,000 jobs/ms, log-normal lengths (mu 6, sigma 1) capped at 8,192 steps, "
                    "+16 KiB per live job per step, chain freed at completion, 64 GiB heap, admission within "
                    "90%% of the heap, %g ms, seed 4" % duration_ms,
        "admission": policy,
        "ops_per_s": st["ops_per_s"],
        "goodput_jobs_per_s": st["goodput_per_s"],
        "run_ms": st["run_ms"],
        "steps": st["steps"],
        "jobs_arrived": st["jobs"],
        "jobs_admitted": st["admitted"],
        "jobs_done": st["done"],
        "jobs_preempted": st["preempted"],
        "preempted_over_admitted": st["preempt_frac"],
        "mean_live_jobs": st["mean_live"],
        "mallocs": st["mallocs"],
        "frees": st["frees"],
        "stalls": st["stalls"],
        "malloc_latency_ns": {"p50": st["lat_p50_ns"], "p99": st["lat_p99_ns"], "max": st["lat_max_ns"]},
        "admission_wait_us": {"p50": st["wait_p50_us"], "p99": st["wait_p99_us"]},
        "phase_us_per_step": st["phase_us"],
        "busy_frac": st["busy_frac"],
        "peak_used_frac": st["peak_used_frac"],
        "peak_projected_frac": st["peak_proj_frac"],
        "projection_checked_steps": st["proj_checked"],
        "projection_mismatches": st["proj_mismatch"],
        "violations": int(st["bad"] + st["aborted"] + st["proj_mismatch"]),
        "work_groups": st["nwg"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--only", choices=["plain", "fused", "batched", "stream", "serve"],
                    help="run only this form of the full-occupancy workload, or (stream) the "
                         "persistent form of the batched stream at the autotuner's usual pick, "
                         "then the copy probe (the commands the counter profiles are taken with); "
                         "(serve) the serving loop with projected admission")
    ap.add_argument("--serve-ms", type=float, default=10000.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every heap is device-local and this build runs on "
              "one MI355X; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import kma_amd

    if kma_amd.lib().kma_device_ok() != 0:
        raise SystemExit("no GPU: " + kma_amd.last_error())

    K, W = args.steps, args.warmup
    if args.only == "serve":
        print(json.dumps(bench_serve(args.serve_ms, kma_amd.ADMIT_PROJECTED)), flush=True)
        return
    if args.only == "stream":
        with kma_amd.Stream(**STREAM) as s:
            cfg = s.heap_config(STREAM_HEAP, 14, kma_amd.TABLE_POW2, 14)
            with kma_amd.Heap.from_config(cfg) as h:
                st = s.run(h, form=kma_amd.FORM_BATCHED, nwg=1024, per_thread=1, warmup=1, reps=K,
                           variant=kma_amd.VAR_AGG_WG8)
        gbps, cms = kma_amd.copy_probe(COPY_BYTES, warmup=1, reps=1)
        print(json.dumps({"form": "stream persistent (8-lane work-group staging, 1,024 x 256, 1 slice per chunk)",
                          "replays": K, "ms_per_replay": st["gpu_ms"], "kernel_ms_avg": st["kernel_ms_avg"],
                          "ops": st["ops"], "nulls": st["nulls"], "bad": st["bad"], "copy_bytes": COPY_BYTES,
                          "copy_gbps": gbps, "copy_ms": cms}), flush=True)
        return
    if args.only:
        form = {"plain": kma_amd.FORM_PLAIN, "fused": kma_amd.FORM_FUSED, "batched": kma_amd.FORM_BATCHED}[args.only]
        with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, SB_LOG2, SBLOCKS, nctx=NCTX) as h:
            st = h.run(kma_amd.WL_LOGUNI, NWI, K, form=form, block=BLOCK, seed=SEED, warmup=W)
        gbps, cms = kma_amd.copy_probe(COPY_BYTES, warmup=1, reps=1)
        print(json.dumps({"form": args.only, "rounds": K, "launches": st["launches"],
                          "kernel_ms_avg": st["kernel_ms_avg"], "ops": st["ops"], "bad": st["bad"],
                          "copy_bytes": COPY_BYTES, "copy_gbps": gbps, "copy_ms": cms}), flush=True)
        return
    res = {}
    for form, name in ((kma_amd.FORM_PLAIN, "plain"), (kma_amd.FORM_FUSED, "fused"),
                       (kma_amd.FORM_BATCHED, "batched")):
        with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, SB_LOG2, SBLOCKS, nctx=NCTX) as h:
            log("bench: %s form, %d warmup + %d timed rounds" % (name, W, K))
            st = h.run(kma_amd.WL_LOGUNI, NWI, K, form=form, block=BLOCK, seed=SEED, warmup=W)
            st["ops_per_s"] = st["ops"] / (st["gpu_ms"] * 1e-3)
            res[name] = st
            log("bench: %s %.3e ops/s, %.3f ms/round, nulls %d bad %d dbl %d badfree %d pool_empty %d waits %s"
                % (name, st["ops_per_s"], st["gpu_ms"] / K, st["nulls"], st["bad"], st["double_frees"],
                   st["bad_frees"], st["pool_empty"], st["waits"]))

    # malloc latency and phase breakdown: separate instrumented runs, so the
    # timed runs above carry no probes
    lat, phases = {}, {}
    for form, name in ((kma_amd.FORM_PLAIN, "plain"), (kma_amd.FORM_BATCHED, "batched")):
        with kma_amd.Heap(kma_amd.LAYOUT_NATIVE, SB_LOG2, SBLOCKS, nctx=NCTX) as h:
            st = h.run(kma_amd.WL_LOGUNI, NWI, 5, form=form, block=BLOCK, seed=SEED, warmup=W,
                       variant=kma_amd.VAR_LATENCY)
            lat[name] = {"p50_ns": st["lat_p50_ns"], "p99_ns": st["lat_p99_ns"], "max_ns": st["lat_max_ns"]}
            st = h.run(kma_amd.WL_LOGUNI, NWI, 5, form=form, block=BLOCK, seed=SEED, round0=100,
                       variant=kma_amd.VAR_PHASES)
            phases[name] = dict(zip(("malloc_ns", "hold_ns", "free_ns"), st["phase_ns"]))
            log("bench: %s malloc latency p50 %.0f ns p99 %.0f ns max %.0f ns; phases per wave-round %s"
                % (name, lat[name]["p50_ns"], lat[name]["p99_ns"], lat[name]["max_ns"],
                   {k: round(v) for k, v in phases[name].items()}))

    peak = {}
    for pat, name in ((0, "distinct"), (2, "same_word_per_wave")):
        ops, _ = kma_amd.atomic_peak(1 << 30, NWI, 64, pat)
        peak[name] = ops
        log("bench: atomicOr peak (%s) %.3e /s" % (name, ops))
    store_peak, _ = kma_amd.atomic_peak(1 << 30, NWI, 64, 3)
    log("bench: scattered 4-B store peak %.3e /s" % store_peak)

    b, p = res["batched"], res["plain"]
    value = b["ops_per_s"]
    dom = b  # dominant kernel: the batched launch
    ops_per_launch = dom["ops"] / max(1, dom["launches"])
    achieved = ops_per_launch * ALGO_BYTES_PER_OP / (dom["kernel_ms_avg"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "heap ops/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": b["gpu_ms"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64 integer",
        "data": "synthetic (device-generated splitmix64 trace, seed 1)",
        "config": {
            "workload": "full occupancy: 1 GiB heap, 16 KiB superblocks, 4,096 allocation contexts, "
                        "524,288 work-items (2,048 x 256), sizes 2^U{3..11} B, hold U{0..64} u32 writes, seed 1",
            "heap_bytes": 1 << 30,
            "contexts": NCTX,
            "work_items": NWI,
            "rounds_timed": K,
            "form": "batched persistent (one launch, all rounds)",
        },
        "plain": {"value": p["ops_per_s"], "ms_per_step": p["gpu_ms"] / K, "launches": p["launches"]},
        "fused_per_lane": {"value": res["fused"]["ops_per_s"], "ms_per_step": res["fused"]["gpu_ms"] / K},
        "batched_over_plain": value / p["ops_per_s"],
        "malloc_latency": lat,
        "phases_per_wave_round": phases,
        "invariants": {k: int(sum(r[k] for r in res.values())) for k in ("nulls", "bad", "double_frees", "bad_frees")},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (traffic_per_round() * K) if traffic_per_round() else None,
        },
        "l2_request_roofline": l2_roofline(b["gpu_ms"] / K, store_peak),
        "atomic_roofline": {
            "peak_distinct_atomics_per_s": peak["distinct"],
            "peak_same_word_per_wave_atomics_per_s": peak["same_word_per_wave"],
            "achieved_ops_per_s": value,
            "frac": value / peak["distinct"],
        },
    }
    if not args.no_stream:
        out["batched_stream"] = bench_stream(5)
    if not args.no_serve:
        out["serving"] = bench_serve(args.serve_ms, kma_amd.ADMIT_PROJECTED)
        out["serving_watermark_admission"] = bench_serve(args.serve_ms, kma_amd.ADMIT_WATERMARK)
    if not args.no_graph:
        out["graph_build"] = bench_graph()
    if not args.no_cpu:
        log("bench: CPU baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
