"""kma_amd - Python view of libkma.so, the MI355X device allocator.

Thin ctypes binding over the C ABI declared in include/kma.h and
include/kma_work.h.  No torch: device memory and streams are owned by the
library.  Loading fails loudly if libkma.so has not been built; there is no
CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkma.so")

LAYOUT_REFERENCE = 0
LAYOUT_NATIVE = 1
MAX_CLASSES = 24

WL_FIXED, WL_LOGUNI, WL_REF_TEST, WL_LOWVAR, WL_HIGHVAR = range(5)
FORM_PLAIN, FORM_BATCHED, FORM_FUSED = 0, 1, 2
VAR_LANE_MALLOC, VAR_LANE_FREE, VAR_LANE_CLAIM = 0x1, 0x2, 0x4
VAR_LATENCY, VAR_PHASES, VAR_LOG = 0x100, 0x200, 0x400

# Exported symbols, as declared in include/*.h (checked by the CPU tests).
EXPORTS = [
    "kma_create", "kma_create_ex", "kma_destroy", "kma_device_heap",
    "kma_heap_bytes", "kma_get_config", "kma_snapshot", "kma_reset",
    "kma_last_error", "kma_device_ok", "kma_run_workload", "kma_trace_serial",
    "kma_trim", "kma_atomic_peak", "kma_check_wave_claim", "kma_wave_free", "kma_copy_probe",
    "kma_stream_create", "kma_stream_destroy", "kma_stream_get_info",
    "kma_stream_trace", "kma_stream_results", "kma_stream_times", "kma_stream_run",
    "kma_stream_heap_config", "kma_stream_autotune",
    "kma_pma_create", "kma_pma_destroy", "kma_pma_device", "kma_pma_reset", "kma_pma_head",
    "kma_pma_run", "kma_queue_create", "kma_queue_destroy", "kma_queue_device", "kma_queue_reset",
    "kma_queue_items", "kma_queue_run", "kma_list_create", "kma_list_destroy", "kma_list_device",
    "kma_list_run_grow", "kma_list_run_get", "kma_list_clear", "kma_list_pages", "kma_tree_run",
    "kma_batch_malloc_det", "kma_batch_free_det", "kma_serve_run", "kma_serve_trace", "kma_serve_lengths",
    "kma_graph_generate", "kma_graph_build", "kma_image_check", "kma_heap_check",
    "kma_batch_malloc_det_noalloc", "kma_det_noalloc_scratch_bytes",
    "kma_lqueue_create", "kma_lqueue_destroy", "kma_lqueue_device", "kma_lqueue_reset", "kma_lqueue_run",
]
DET_NOALLOC_MAX = 524288
GRAPH_PLAIN, GRAPH_CACHE, GRAPH_COLLECTIVE = 0, 1, 2
TABLE_POW2, TABLE_HALF = 0, 1
TUNE_MAX = 128
VAR_AGG_WAVE, VAR_AGG_WG8, VAR_AGG_WG16, VAR_AGG_WG32 = 0x00, 0x10, 0x20, 0x30


class Config(ctypes.Structure):
    _fields_ = [
        ("layout", ctypes.c_uint32),
        ("sb_log2", ctypes.c_uint32),
        ("sblocks", ctypes.c_uint64),
        ("nctx", ctypes.c_uint32),
        ("npool", ctypes.c_uint32),
        ("nclasses", ctypes.c_uint32),
        ("class_size", ctypes.c_uint32 * MAX_CLASSES),
        ("page_log2", ctypes.c_uint32),
        ("npages", ctypes.c_uint64),
        ("no_partial", ctypes.c_uint32),
    ]


class Workload(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_uint32),
        ("form", ctypes.c_uint32),
        ("nwi", ctypes.c_uint64),
        ("block", ctypes.c_uint32),
        ("rounds", ctypes.c_uint32),
        ("round0", ctypes.c_uint32),
        ("fixed", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("warmup", ctypes.c_uint32),
        ("deadline_ms", ctypes.c_uint32),
        ("variant", ctypes.c_uint32),
    ]


class RunStats(ctypes.Structure):
    _fields_ = [
        ("gpu_ms", ctypes.c_double),
        ("kernel_ms_avg", ctypes.c_double),
        ("launches", ctypes.c_uint64),
        ("ops", ctypes.c_uint64),
        ("mallocs", ctypes.c_uint64),
        ("nulls", ctypes.c_uint64),
        ("bad", ctypes.c_uint64),
        ("double_frees", ctypes.c_uint64),
        ("bad_frees", ctypes.c_uint64),
        ("pool_empty", ctypes.c_uint64),
        ("waits", ctypes.c_uint64 * 4),
        ("remote_frees", ctypes.c_uint64),
        ("migrated", ctypes.c_uint64),
        ("lat_p50_ns", ctypes.c_double),
        ("lat_p99_ns", ctypes.c_double),
        ("lat_max_ns", ctypes.c_double),
        ("phase_ns", ctypes.c_double * 3),
        ("deadline_hits", ctypes.c_uint64),
        ("listed", ctypes.c_uint64),
        ("reattached", ctypes.c_uint64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["waits"] = list(self.waits)
        d["phase_ns"] = list(self.phase_ns)
        return d


class StreamSpec(ctypes.Structure):
    _fields_ = [
        ("nbatch", ctypes.c_uint32),
        ("max_batch", ctypes.c_uint32),
        ("zipf_milli", ctypes.c_uint32),
        ("free_pct", ctypes.c_uint32),
        ("min_size", ctypes.c_uint32),
        ("max_size", ctypes.c_uint32),
        ("page_min", ctypes.c_uint32),
        ("drain", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
    ]


class StreamInfo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in (
        "nreq", "nbatch", "nmalloc", "nfree", "max_batch", "peak_live", "small_bytes",
        "large_bytes", "peak_small_bytes", "peak_large_bytes")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StreamStats(ctypes.Structure):
    _fields_ = [
        ("gpu_ms", ctypes.c_double),
        ("kernel_ms_avg", ctypes.c_double),
    ] + [(k, ctypes.c_uint64) for k in (
        "launches", "ops", "mallocs", "nulls", "frees", "free_nulls", "bad", "wait_timeouts",
        "large", "large_nulls", "double_frees", "bad_frees", "pool_empty")] + [
        ("waits", ctypes.c_uint64 * 4),
        ("remote_frees", ctypes.c_uint64),
        ("tick_ns", ctypes.c_double),
        ("phase_ns", ctypes.c_double * 3),
        ("listed", ctypes.c_uint64),
        ("reattached", ctypes.c_uint64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["waits"] = list(self.waits)
        d["phase_ns"] = list(self.phase_ns)
        return d


class TuneTrial(ctypes.Structure):
    _fields_ = [
        ("sb_log2", ctypes.c_uint32), ("table", ctypes.c_uint32), ("nwg", ctypes.c_uint32),
        ("per_thread", ctypes.c_uint32), ("agg", ctypes.c_uint32), ("no_partial", ctypes.c_uint32),
        ("ops_per_s", ctypes.c_double),
        ("ms", ctypes.c_double),
        ("nulls", ctypes.c_uint64), ("violations", ctypes.c_uint64),
    ]


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("cfg", Config), ("nwg", ctypes.c_uint32), ("per_thread", ctypes.c_uint32),
        ("table", ctypes.c_uint32), ("ntrials", ctypes.c_uint32), ("agg", ctypes.c_uint32),
        ("ops_per_s", ctypes.c_double),
        ("nulls", ctypes.c_uint64), ("trials", TuneTrial * TUNE_MAX),
    ]


def config_dict(c):
    return {"layout": c.layout, "sb_log2": c.sb_log2, "sblocks": c.sblocks, "nctx": c.nctx,
            "npool": c.npool, "classes": list(c.class_size[:c.nclasses]), "page_log2": c.page_log2,
            "npages": c.npages, "no_partial": c.no_partial}


class ServeSpec(ctypes.Structure):
    _fields_ = [
        ("rate_per_ms", ctypes.c_double), ("len_mu", ctypes.c_double), ("len_sigma", ctypes.c_double),
        ("duration_ms", ctypes.c_double), ("max_len", ctypes.c_uint32), ("block_bytes", ctypes.c_uint32),
        ("watermark_pct", ctypes.c_uint32), ("max_live", ctypes.c_uint32), ("max_admit", ctypes.c_uint32),
        ("nwg", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("admit", ctypes.c_uint32), ("keep_live", ctypes.c_uint32),
    ]


class GraphStats(ctypes.Structure):
    _fields_ = [("gpu_ms", ctypes.c_double)] + [
        (k, ctypes.c_uint64) for k in ("edges", "nodes", "node_mallocs", "node_frees", "link_mallocs", "nulls",
                                       "bad", "violations", "links_found")] + [
        ("nwg", ctypes.c_uint32), ("pad", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class HeapReport(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in (
        "problems", "sb_total", "sb_free", "sb_in_spares", "sb_heads", "sb_detached", "orphans", "live_blocks",
        "live_bytes", "page_runs", "pages_used", "pages_total", "sb_unqueued", "sb_listed")] + [
        ("counters", ctypes.c_uint64 * 11), ("first", ctypes.c_char * 1024)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_[:14]}
        d["counters"] = list(self.counters)
        d["first"] = self.first.decode(errors="replace").splitlines()
        return d


class ServeStats(ctypes.Structure):
    _fields_ = [("run_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double)] + [
        (k, ctypes.c_uint64) for k in ("jobs", "truncated", "steps", "admitted", "done", "drained", "mallocs",
                                       "frees", "stalls", "drain_frees", "bad", "preempted")] + [
        (k, ctypes.c_double) for k in ("ops_per_s", "lat_p50_ns", "lat_p99_ns", "lat_max_ns", "wait_p50_us",
                                       "wait_p99_us", "peak_used_frac")] + [
        ("nwg", ctypes.c_uint32), ("aborted", ctypes.c_uint32), ("admit", ctypes.c_uint32), ("pad", ctypes.c_uint32),
        ("cap_blocks", ctypes.c_uint64), ("rebuilds", ctypes.c_uint64), ("proj_checked", ctypes.c_uint64),
        ("proj_mismatch", ctypes.c_uint64)] + [
        (k, ctypes.c_double) for k in ("goodput_per_s", "preempt_frac", "mean_live", "peak_proj_frac")] + [
        ("phase_us", ctypes.c_double * 4), ("busy_frac", ctypes.c_double * 2)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pad", "phase_us", "busy_frac")}
        d["phase_us"] = dict(zip(("admit_work", "admit", "grow", "free"), self.phase_us))
        d["busy_frac"] = dict(zip(("grow", "free"), self.busy_frac))
        return d


ADMIT_PROJECTED = 0
ADMIT_WATERMARK = 1


def serve_spec(rate_per_ms=2000.0, len_mu=6.0, len_sigma=1.0, duration_ms=10000.0, max_len=8192,
               block_bytes=16384, watermark_pct=90, max_live=32768, max_admit=0, nwg=0, seed=4,
               admit=ADMIT_PROJECTED, keep_live=0):
    return ServeSpec(rate_per_ms, len_mu, len_sigma, duration_ms, max_len, block_bytes, watermark_pct, max_live,
                     max_admit, nwg, seed, admit, keep_live)


class KmaError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libkma.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KmaError("libkma.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        P = ctypes.POINTER
        L.kma_create.argtypes = [ctypes.c_uint]
        L.kma_create.restype = vp
        L.kma_create_ex.argtypes = [P(Config)]
        L.kma_create_ex.restype = vp
        L.kma_destroy.argtypes = [vp]
        L.kma_destroy.restype = None
        L.kma_device_heap.argtypes = [vp]
        L.kma_device_heap.restype = vp
        L.kma_heap_bytes.argtypes = [vp]
        L.kma_heap_bytes.restype = u64
        L.kma_get_config.argtypes = [vp, P(Config)]
        L.kma_snapshot.argtypes = [vp, u64, vp, u64]
        L.kma_reset.argtypes = [vp]
        L.kma_last_error.restype = ctypes.c_char_p
        L.kma_device_ok.restype = ctypes.c_int
        L.kma_run_workload.argtypes = [vp, P(Workload), P(RunStats)]
        L.kma_trace_serial.argtypes = [vp, vp, u64, vp, vp]
        L.kma_trim.argtypes = [vp]
        L.kma_check_wave_claim.argtypes = [vp, u32, vp, vp, vp, vp, vp, vp, vp]
        L.kma_wave_free.argtypes = [vp, vp, u32]
        L.kma_atomic_peak.argtypes = [u64, u64, u32, ctypes.c_int, u32, u32,
                                      P(ctypes.c_double), P(ctypes.c_double)]
        L.kma_copy_probe.argtypes = [u64, u32, u32, P(ctypes.c_double), P(ctypes.c_double)]
        L.kma_stream_create.argtypes = [P(StreamSpec)]
        L.kma_stream_create.restype = vp
        L.kma_stream_destroy.argtypes = [vp]
        L.kma_stream_destroy.restype = None
        L.kma_stream_get_info.argtypes = [vp, P(StreamInfo)]
        L.kma_stream_trace.argtypes = [vp, vp, vp]
        L.kma_stream_results.argtypes = [vp, vp]
        L.kma_stream_times.argtypes = [vp, vp]
        L.kma_stream_run.argtypes = [vp, vp, u32, u32, u32, u32, u32, u32, P(StreamStats)]
        L.kma_stream_heap_config.argtypes = [vp, u64, u32, u32, u32, P(Config)]
        i64 = ctypes.c_int64
        L.kma_pma_create.argtypes = [u64, u32]
        L.kma_pma_create.restype = vp
        L.kma_pma_destroy.argtypes = [vp]
        L.kma_pma_destroy.restype = None
        L.kma_pma_device.argtypes = [vp]
        L.kma_pma_device.restype = vp
        L.kma_pma_reset.argtypes = [vp]
        L.kma_pma_head.argtypes = [vp]
        L.kma_pma_head.restype = u64
        L.kma_pma_run.argtypes = [vp, vp, u64, ctypes.c_int, vp]
        L.kma_queue_create.argtypes = [u64]
        L.kma_queue_create.restype = vp
        L.kma_queue_destroy.argtypes = [vp]
        L.kma_queue_destroy.restype = None
        L.kma_queue_device.argtypes = [vp]
        L.kma_queue_device.restype = vp
        L.kma_queue_reset.argtypes = [vp]
        L.kma_queue_items.argtypes = [vp, vp, u64]
        L.kma_queue_items.restype = i64
        L.kma_queue_run.argtypes = [vp, u32, u64, u32, P(u64), vp]
        L.kma_lqueue_create.argtypes = []
        L.kma_lqueue_create.restype = vp
        L.kma_lqueue_destroy.argtypes = [vp]
        L.kma_lqueue_destroy.restype = None
        L.kma_lqueue_device.argtypes = [vp]
        L.kma_lqueue_device.restype = vp
        L.kma_lqueue_reset.argtypes = [vp]
        L.kma_lqueue_run.argtypes = [vp, u32, u64, u32, vp, P(LQueueReport)]
        L.kma_list_create.argtypes = [vp, u32, u64]
        L.kma_list_create.restype = vp
        L.kma_list_destroy.argtypes = [vp]
        L.kma_list_destroy.restype = None
        L.kma_list_device.argtypes = [vp]
        L.kma_list_device.restype = vp
        L.kma_list_run_grow.argtypes = [vp, vp, u64, u32, ctypes.c_int, vp]
        L.kma_list_run_get.argtypes = [vp, vp, u64, vp]
        L.kma_list_clear.argtypes = [vp, u64]
        L.kma_list_pages.argtypes = [vp, vp, u64]
        L.kma_list_pages.restype = i64
        L.kma_tree_run.argtypes = [vp, vp, u64, vp, vp, u64, vp, vp, P(u64), P(u64)]
        L.kma_batch_malloc_det.argtypes = [vp, vp, u64, vp, P(ctypes.c_double)]
        L.kma_batch_malloc_det.restype = i64
        L.kma_batch_free_det.argtypes = [vp, vp, u64, P(ctypes.c_double)]
        L.kma_batch_malloc_det_noalloc.argtypes = [vp, vp, u64, vp, vp, vp]
        L.kma_det_noalloc_scratch_bytes.argtypes = []
        L.kma_det_noalloc_scratch_bytes.restype = u64
        L.kma_serve_run.argtypes = [vp, P(ServeSpec), vp, u64, vp, u64, P(ServeStats)]
        L.kma_serve_trace.argtypes = [P(ServeSpec), vp, vp, u64]
        L.kma_serve_trace.restype = i64
        L.kma_serve_lengths.argtypes = [P(ServeSpec), vp, u64, vp]
        L.kma_graph_generate.argtypes = [u32, u64, u64, vp]
        L.kma_graph_build.argtypes = [vp, vp, vp, u64, ctypes.c_int, u32, ctypes.c_int, P(GraphStats)]
        L.kma_image_check.argtypes = [vp, u64, P(HeapReport)]
        L.kma_heap_check.argtypes = [vp, P(HeapReport)]
        L.kma_stream_autotune.argtypes = [vp, u64, u32, u64, u32, P(TuneResult)]
        _lib = L
    return _lib


def last_error():
    return lib().kma_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise KmaError("%s failed: %s" % (what, last_error()))


class Heap:
    """A device heap (kma_create / kma_create_ex)."""

    def __init__(self, layout=LAYOUT_REFERENCE, sb_log2=12, sblocks=4096, nctx=0,
                 npool=0, classes=None, page_log2=0, npages=0, no_partial=0):
        c = Config()
        c.layout = layout
        c.sb_log2 = sb_log2
        c.sblocks = sblocks
        c.nctx = nctx
        c.npool = npool
        if classes:
            c.nclasses = len(classes)
            for i, s in enumerate(classes):
                c.class_size[i] = s
        c.page_log2 = page_log2
        c.npages = npages
        c.no_partial = no_partial
        self._h = lib().kma_create_ex(ctypes.byref(c))
        if not self._h:
            raise KmaError("kma_create_ex: " + last_error())
        self.layout = layout

    @classmethod
    def from_config(cls, c):
        """kma_create_ex on a Config structure."""
        self = cls.__new__(cls)
        self._h = lib().kma_create_ex(ctypes.byref(c))
        if not self._h:
            raise KmaError("kma_create_ex: " + last_error())
        self.layout = c.layout
        return self

    @classmethod
    def create(cls, sblocks):
        """Reference-compatible constructor: kma_create(sblocks)."""
        self = cls.__new__(cls)
        self._h = lib().kma_create(sblocks)
        if not self._h:
            raise KmaError("kma_create: " + last_error())
        self.layout = LAYOUT_REFERENCE
        return self

    def close(self):
        if getattr(self, "_h", None):
            lib().kma_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def device_ptr(self):
        return lib().kma_device_heap(self._h)

    @property
    def nbytes(self):
        return lib().kma_heap_bytes(self._h)

    def config(self):
        c = Config()
        _check(lib().kma_get_config(self._h, ctypes.byref(c)), "kma_get_config")
        return c

    def snapshot(self, offset=0, nbytes=None):
        if nbytes is None:
            nbytes = self.nbytes - offset
        buf = np.empty(nbytes, dtype=np.uint8)
        _check(lib().kma_snapshot(self._h, offset, buf.ctypes.data, nbytes), "kma_snapshot")
        return buf

    def reset(self):
        _check(lib().kma_reset(self._h), "kma_reset")

    def check(self):
        """Walk a quiescent native heap (kma_heap_check): a dict whose
        `problems` is 0 when every superblock, shard and page run is
        consistent; `first` lists the first problems found."""
        r = HeapReport()
        _check(lib().kma_heap_check(self._h, ctypes.byref(r)), "kma_heap_check")
        return r.as_dict()

    def trim(self):
        _check(lib().kma_trim(self._h), "kma_trim")

    def trace_serial(self, ops, pat=None):
        ops = np.ascontiguousarray(ops, dtype=np.int64)
        out = np.zeros(len(ops), dtype=np.uint64)
        patp = None
        if pat is not None:
            pat = np.ascontiguousarray(pat, dtype=np.uint32)
            patp = pat.ctypes.data
        _check(lib().kma_trace_serial(self._h, ops.ctypes.data, len(ops), out.ctypes.data, patp),
               "kma_trace_serial")
        return out

    def wave_free(self, offs):
        """One wavefront frees the blocks at these heap offsets (<= 64, lane i
        frees offs[i], 0 = none) in a single kma_free call (test hook)."""
        offs = np.ascontiguousarray(offs, dtype=np.uint64)
        _check(lib().kma_wave_free(self._h, offs.ctypes.data, len(offs)), "kma_wave_free")

    def batch_malloc_det(self, sizes):
        """Deterministic batch malloc (reference layout): (offsets, NULL count,
        device ms)."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        out = np.zeros(len(sizes), dtype=np.uint64)
        ms = ctypes.c_double()
        nulls = lib().kma_batch_malloc_det(self._h, sizes.ctypes.data, len(sizes), out.ctypes.data,
                                           ctypes.byref(ms))
        if nulls < 0:
            raise KmaError("kma_batch_malloc_det: " + last_error())
        return out, nulls, ms.value

    def batch_malloc_det_noalloc(self, d_sizes, n, d_out, d_status, stream=None):
        """Allocation-free deterministic batch (kma_batch_malloc_det_noalloc):
        every argument is a raw device address (int) owned by the caller; the
        kernels are enqueued on `stream` (a hipStream_t as int, None = null
        stream) and the call returns without waiting."""
        _check(lib().kma_batch_malloc_det_noalloc(self._h, d_sizes, n, d_out, d_status, stream),
               "kma_batch_malloc_det_noalloc")

    def batch_free_det(self, offs):
        """Free the preceding deterministic batch in index order; device ms."""
        offs = np.ascontiguousarray(offs, dtype=np.uint64)
        ms = ctypes.c_double()
        _check(lib().kma_batch_free_det(self._h, offs.ctypes.data, len(offs), ctypes.byref(ms)),
               "kma_batch_free_det")
        return ms.value

    def run(self, mode, nwi, rounds, form=FORM_PLAIN, block=64, fixed=64, seed=1,
            round0=0, warmup=0, variant=0, deadline_ms=0):
        w = Workload(mode, form, nwi, block, rounds, round0, fixed, seed, warmup, deadline_ms, variant)
        st = RunStats()
        _check(lib().kma_run_workload(self._h, ctypes.byref(w), ctypes.byref(st)), "kma_run_workload")
        return st.as_dict()


def check_wave_claim(heap, cls, base, got, bm_in):
    """Run the wavefront bitmap claim on preset bitmaps (test hook)."""
    n = len(cls)
    cls = np.ascontiguousarray(cls, dtype=np.uint32)
    base = np.ascontiguousarray(base, dtype=np.uint32)
    got = np.ascontiguousarray(got, dtype=np.uint32)
    bm_in = np.ascontiguousarray(bm_in, dtype=np.uint32)
    res = np.zeros(n * 64, dtype=np.uint64)
    bm_out = np.zeros_like(bm_in)
    done = np.zeros(n, dtype=np.uint32)
    _check(lib().kma_check_wave_claim(heap._h, n, cls.ctypes.data, base.ctypes.data, got.ctypes.data,
                                      bm_in.ctypes.data, res.ctypes.data, bm_out.ctypes.data,
                                      done.ctypes.data), "kma_check_wave_claim")
    return res.reshape(n, 64), bm_out, done


def atomic_peak(nbytes, nthreads, iters, pattern, warmup=2, reps=5):
    ops = ctypes.c_double()
    ms = ctypes.c_double()
    _check(lib().kma_atomic_peak(nbytes, nthreads, iters, pattern, warmup, reps, ctypes.byref(ops),
                                 ctypes.byref(ms)), "kma_atomic_peak")
    return ops.value, ms.value


class Stream:
    """A batched request stream (kma_stream_create): Zipf batch sizes,
    log-uniform request sizes, frees of earlier batches' blocks, and an
    optional drain batch that frees everything still live."""

    def __init__(self, nbatch, max_batch, zipf=1.2, free_pct=40, min_size=16, max_size=65536,
                 page_min=16384, drain=True, seed=3):
        sp = StreamSpec(nbatch, max_batch, int(round(zipf * 1000)), free_pct, min_size, max_size,
                        page_min, 1 if drain else 0, seed)
        self._s = lib().kma_stream_create(ctypes.byref(sp))
        if not self._s:
            raise KmaError("kma_stream_create: " + last_error())
        inf = StreamInfo()
        _check(lib().kma_stream_get_info(self._s, ctypes.byref(inf)), "kma_stream_get_info")
        self.info = inf.as_dict()

    def close(self):
        if getattr(self, "_s", None):
            lib().kma_stream_destroy(self._s)
            self._s = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def trace(self):
        """(req u32[nreq], batch_off u32[nbatch + 1])"""
        req = np.zeros(self.info["nreq"], dtype=np.uint32)
        boff = np.zeros(self.info["nbatch"] + 1, dtype=np.uint32)
        _check(lib().kma_stream_trace(self._s, req.ctypes.data, boff.ctypes.data), "kma_stream_trace")
        return req, boff

    def results(self):
        res = np.zeros(self.info["nreq"], dtype=np.uint64)
        _check(lib().kma_stream_results(self._s, res.ctypes.data), "kma_stream_results")
        return res

    def times(self):
        t = np.zeros(self.info["nreq"], dtype=np.uint64)
        _check(lib().kma_stream_times(self._s, t.ctypes.data), "kma_stream_times")
        return t

    def run(self, heap, form=FORM_BATCHED, nwg=1024, per_thread=4, warmup=0, reps=1, variant=0):
        st = StreamStats()
        _check(lib().kma_stream_run(heap._h, self._s, form, nwg, per_thread, warmup, reps, variant,
                                    ctypes.byref(st)), "kma_stream_run")
        return st.as_dict()

    def heap_config(self, heap_bytes, sb_log2=14, table=TABLE_POW2, page_log2=14):
        """Native heap Config for this stream (kma_stream_heap_config)."""
        c = Config()
        _check(lib().kma_stream_heap_config(self._s, heap_bytes, sb_log2, table, page_log2,
                                            ctypes.byref(c)), "kma_stream_heap_config")
        return c

    def autotune(self, heap_bytes, page_log2=14, null_limit=0, reps=2):
        """Sweep superblock size, class table and launch shape for the
        persistent form (kma_stream_autotune)."""
        r = TuneResult()
        _check(lib().kma_stream_autotune(self._s, heap_bytes, page_log2, null_limit, reps,
                                         ctypes.byref(r)), "kma_stream_autotune")
        trials = [{k: getattr(t, k) for k, _ in TuneTrial._fields_} for t in r.trials[:r.ntrials]]
        return {"cfg": r.cfg, "config": config_dict(r.cfg), "nwg": r.nwg, "per_thread": r.per_thread,
                "table": r.table, "agg": r.agg, "ops_per_s": r.ops_per_s, "nulls": r.nulls, "trials": trials}


def splitmix(seed, gid, k):
    """splitmix64 finaliser of seed*K1 + gid*K2 + k*K3 (vectorised over gid)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + np.asarray(gid, dtype=np.uint64)
             * np.uint64(0xBF58476D1CE4E5B9) + np.uint64(k) * np.uint64(0x94D049BB133111EB)) & M
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def loguniform_sizes(n, lo_log2, hi_log2, seed):
    """n sizes log-uniform over [2^lo_log2, 2^hi_log2): exponent e uniform,
    then uniform within the octave, integer only."""
    gid = np.arange(n, dtype=np.uint64)
    e = np.uint64(lo_log2) + splitmix(seed, gid, 0) % np.uint64(hi_log2 - lo_log2)
    m = (splitmix(seed, gid, 1) >> np.uint64(16)) % (np.uint64(1) << e)
    return ((np.uint64(1) << e) + m).astype(np.uint32)


class _Handle:
    _destroy = None

    def close(self):
        if getattr(self, "_p", None):
            getattr(lib(), self._destroy)(self._p)
            self._p = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Pma(_Handle):
    """Bump allocator (kma_pma_create; the reference's pma_create)."""
    _destroy = "kma_pma_destroy"

    def __init__(self, nbytes, align=1):
        self._p = lib().kma_pma_create(nbytes, align)
        if not self._p:
            raise KmaError("kma_pma_create: " + last_error())

    def reset(self):
        _check(lib().kma_pma_reset(self._p), "kma_pma_reset")

    @property
    def head(self):
        return lib().kma_pma_head(self._p)

    def run(self, sizes, serial=False):
        """Data offsets of the blocks (UINT64_MAX for NULL)."""
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        out = np.zeros(len(sizes), dtype=np.uint64)
        _check(lib().kma_pma_run(self._p, sizes.ctypes.data, len(sizes), 1 if serial else 0, out.ctypes.data),
               "kma_pma_run")
        return out


class Queue(_Handle):
    """MPMC pointer queue (kma_queue_create; the reference's clQueue_create)."""
    _destroy = "kma_queue_destroy"

    def __init__(self, capacity):
        self._p = lib().kma_queue_create(capacity)
        if not self._p:
            raise KmaError("kma_queue_create: " + last_error())

    def reset(self):
        _check(lib().kma_queue_reset(self._p), "kma_queue_reset")

    def items(self, maxn=1 << 20):
        buf = np.zeros(maxn, dtype=np.uint64)
        n = lib().kma_queue_items(self._p, buf.ctypes.data, maxn)
        if n < 0:
            raise KmaError("kma_queue_items: " + last_error())
        return buf[:min(n, maxn)]

    def run(self, mode, nwi, rounds=0):
        """(item base address, dequeue order for mode 2 or None)"""
        base = ctypes.c_uint64()
        order = np.zeros(max(nwi, 1), dtype=np.int64) if mode == 2 else None
        _check(lib().kma_queue_run(self._p, mode, nwi, rounds, ctypes.byref(base),
                                   order.ctypes.data if order is not None else None), "kma_queue_run")
        return base.value, (order[:nwi] if order is not None else None)


class LQueueReport(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("items", "tail_ok", "cycle", "stubs", "outside", "repeats", "lost")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LinkedQueue(_Handle):
    """Unbounded intrusive MPMC queue (kma_lqueue_create; the reference's
    clQueue_create with its contract: items linked through their first word)."""
    _destroy = "kma_lqueue_destroy"

    def __init__(self):
        self._p = lib().kma_lqueue_create()
        if not self._p:
            raise KmaError("kma_lqueue_create: " + last_error())

    def reset(self):
        _check(lib().kma_lqueue_reset(self._p), "kma_lqueue_reset")

    def run(self, mode, nwi, rounds=0):
        """(report dict, order array for mode 2 or None)."""
        rep = LQueueReport()
        order = np.zeros(max(nwi, 1), dtype=np.int64) if mode == 2 else None
        _check(lib().kma_lqueue_run(self._p, mode, nwi, rounds, order.ctypes.data if order is not None else None,
                                    ctypes.byref(rep)), "kma_lqueue_run")
        return rep.as_dict(), (order[:nwi] if order is not None else None)


class ArrayList(_Handle):
    """Growable array list over a native heap (kma_list_create; the
    reference's clArrayList_create)."""
    _destroy = "kma_list_destroy"

    def __init__(self, heap, obj_size, max_pages=1 << 16):
        self._p = lib().kma_list_create(heap._h, obj_size, max_pages)
        if not self._p:
            raise KmaError("kma_list_create: " + last_error())
        self.obj_size = obj_size

    def grow(self, objs, block=256, local=True):
        objs = np.ascontiguousarray(objs, dtype=np.uint32)
        out = np.zeros(len(objs), dtype=np.uint64)
        _check(lib().kma_list_run_grow(self._p, objs.ctypes.data, len(objs), block, 1 if local else 0,
                                       out.ctypes.data), "kma_list_run_grow")
        return out

    def get(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        out = np.zeros(len(idx), dtype=np.uint64)
        _check(lib().kma_list_run_get(self._p, idx.ctypes.data, len(idx), out.ctypes.data), "kma_list_run_get")
        return out

    def clear(self, nwi=1):
        _check(lib().kma_list_clear(self._p, nwi), "kma_list_clear")

    def pages(self, maxn=1 << 16):
        buf = np.zeros(maxn, dtype=np.uint64)
        n = lib().kma_list_pages(self._p, buf.ctypes.data, maxn)
        if n < 0:
            raise KmaError("kma_list_pages: " + last_error())
        return buf[:min(n, maxn)]


def tree_run(heap, keys, query):
    """Concurrent kma_tree_add of keys, kma_tree_get of query, in-order walk;
    returns (added, found, inorder keys, violations)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    query = np.ascontiguousarray(query, dtype=np.uint32)
    added = np.zeros(len(keys), dtype=np.uint32)
    found = np.zeros(len(query), dtype=np.uint32)
    inorder = np.zeros(len(keys) + 1, dtype=np.uint32)
    n = ctypes.c_uint64()
    bad = ctypes.c_uint64()
    _check(lib().kma_tree_run(heap._h, keys.ctypes.data, len(keys), added.ctypes.data, query.ctypes.data,
                              len(query), found.ctypes.data, inorder.ctypes.data, ctypes.byref(n),
                              ctypes.byref(bad)), "kma_tree_run")
    return added, found, inorder[:min(n.value, len(keys) + 1)], bad.value


def serve_run(heap, spec, jobs=None, nresults=0):
    """Run the serving loop (kma_serve_run): (stats dict, results array)."""
    st = ServeStats()
    jp, nj = None, 0
    if jobs is not None:
        jobs = np.ascontiguousarray(jobs, dtype=np.uint64)
        jp, nj = jobs.ctypes.data, len(jobs)
        nresults = nj
    res = np.zeros(max(nresults, 1), dtype=np.uint64)
    _check(lib().kma_serve_run(heap._h, ctypes.byref(spec), jp, nj, res.ctypes.data if nresults else None,
                               nresults, ctypes.byref(st)), "kma_serve_run")
    return st.as_dict(), res[:nresults]


def serve_trace(spec, maxn=1 << 26):
    """(arrival ticks, lengths) of the serving trace (host only)."""
    n = lib().kma_serve_trace(ctypes.byref(spec), None, None, 0)
    if n < 0:
        raise KmaError("kma_serve_trace: " + last_error())
    n = min(n, maxn)
    arr = np.zeros(max(n, 1), dtype=np.uint64)
    ln = np.zeros(max(n, 1), dtype=np.uint32)
    lib().kma_serve_trace(ctypes.byref(spec), arr.ctypes.data, ln.ctypes.data, n)
    return arr[:n], ln[:n]


def serve_lengths(spec, jobs):
    jobs = np.ascontiguousarray(jobs, dtype=np.uint64)
    out = np.zeros(max(len(jobs), 1), dtype=np.uint32)
    _check(lib().kma_serve_lengths(ctypes.byref(spec), jobs.ctypes.data, len(jobs), out.ctypes.data),
           "kma_serve_lengths")
    return out[:len(jobs)]


def image_check(image):
    """kma_image_check on a metadata image (a uint8 array holding bytes
    [0, off_data) of a native heap)."""
    image = np.ascontiguousarray(image, dtype=np.uint8)
    r = HeapReport()
    _check(lib().kma_image_check(image.ctypes.data, image.size, ctypes.byref(r)), "kma_image_check")
    return r.as_dict()


def graph_generate(nnodes, nedges, seed):
    """An edge list: nedges (source, sink) rows over nnodes distinct ids."""
    out = np.zeros((max(nedges, 1), 2), dtype=np.uint32)
    _check(lib().kma_graph_generate(nnodes, nedges, seed, out.ctypes.data), "kma_graph_generate")
    return out[:nedges]


def graph_build(edges, mode, heap=None, pma=None, nwg=0, teardown=False):
    """Build the graph of `edges` ((n, 2) uint32) on a native heap or a bump
    allocator (kma_graph_build); returns its statistics."""
    edges = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 2)
    st = GraphStats()
    _check(lib().kma_graph_build(heap._h if heap is not None else None, pma._p if pma is not None else None,
                                 edges.ctypes.data, len(edges), mode, nwg, 1 if teardown else 0,
                                 ctypes.byref(st)), "kma_graph_build")
    return st.as_dict()


def graph_write(path, edges):
    """Write an edge list in the reference's data-set format: the number of
    entries, then one `{source,sink}` line per edge (tb_clTree.c:350-389)."""
    edges = np.asarray(edges, dtype=np.uint32).reshape(-1, 2)
    with open(path, "w") as f:
        f.write("%d\n" % len(edges))
        f.writelines("{%d,%d}\n" % (int(a), int(b)) for a, b in edges)


def graph_read(path):
    """Read an edge list in the reference's data-set format."""
    with open(path) as f:
        n = int(f.readline())
        rows = []
        for _ in range(n):
            line = f.readline().strip()
            if not (line.startswith("{") and line.endswith("}")):
                raise ValueError("malformed data-set line: %r" % line)
            a, b = line[1:-1].split(",")
            rows.append((int(a), int(b)))
    return np.array(rows, dtype=np.uint32).reshape(-1, 2)


def serve_pattern(job, k):
    """The 64-bit pattern the serving loop stamps into block k of job `job`
    (vectorised over k)."""
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        z = (np.uint64(job) * np.uint64(0x9E3779B97F4A7C15) + np.asarray(k, dtype=np.uint64)
             * np.uint64(0xBF58476D1CE4E5B9) + np.uint64(0x94D049BB133111EB)) & M
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def copy_probe(nbytes, warmup=2, reps=5):
    """(GB/s read+written, ms per launch) of a streaming device copy."""
    g = ctypes.c_double()
    ms = ctypes.c_double()
    _check(lib().kma_copy_probe(nbytes, warmup, reps, ctypes.byref(g), ctypes.byref(ms)), "kma_copy_probe")
    return g.value, ms.value
