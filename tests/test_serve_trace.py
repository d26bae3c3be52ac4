"""The serving loop's job trace, generated on the host (no device): Poisson
arrivals at the stated rate and log-normal lengths."""
import numpy as np

import kma_amd


def test_trace_distributions():
    sp = kma_amd.serve_spec(rate_per_ms=2000.0, len_mu=6.0, len_sigma=1.0, duration_ms=100.0, seed=4)
    arr, ln = kma_amd.serve_trace(sp)
    tick_per_ms = 1e5  # no device here: the 100 MHz wall-clock default
    n = len(arr)
    assert abs(n - 200000) < 4 * np.sqrt(200000)
    gaps = np.diff(arr.astype(np.float64))
    assert abs(gaps.mean() - tick_per_ms / 2000.0) < 0.5
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.02  # exponential: CV = 1
    assert (np.diff(arr.astype(np.int64)) >= 0).all() and arr[-1] < 100 * tick_per_ms
    lg = np.log(ln[ln < 8192].astype(np.float64))
    assert abs(np.median(np.log(ln)) - 6.0) < 0.02
    assert abs(lg.std() - 1.0) < 0.03
    assert ln.min() >= 1 and ln.max() <= 8192


def test_trace_deterministic_and_lengths_agree():
    sp = kma_amd.serve_spec(rate_per_ms=100.0, duration_ms=50.0, len_mu=4.0, len_sigma=1.0, max_len=512, seed=9)
    a1, l1 = kma_amd.serve_trace(sp)
    a2, l2 = kma_amd.serve_trace(sp)
    assert np.array_equal(a1, a2) and np.array_equal(l1, l2)
    jobs = np.arange(0, len(l1), 7, dtype=np.uint64)
    assert np.array_equal(kma_amd.serve_lengths(sp, jobs), l1[jobs.astype(np.int64)])
    assert l1.max() <= 512
