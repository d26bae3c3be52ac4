"""ctypes view of the oracle (oracle/refmodel.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the bench's
CPU-baseline leg as the checker.  The product path (kma_amd, libkma.so) never
imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "librefmodel.so")

RM_OK, RM_HANG, RM_EINVAL = 0, 1, 2


class RmHeap(ctypes.Structure):
    _fields_ = [
        ("img", ctypes.POINTER(ctypes.c_uint8)),
        ("bytes", ctypes.c_uint64),
        ("base", ctypes.c_uint64),
        ("log2", ctypes.c_uint32),
        ("sb", ctypes.c_uint32),
        ("buckets", ctypes.c_uint32),
        ("hdr", ctypes.c_uint32),
        ("nsb", ctypes.c_uint32),
        ("owns_img", ctypes.c_int),
    ]


class RmtResult(ctypes.Structure):
    _fields_ = [
        ("seconds", ctypes.c_double),
        ("ops", ctypes.c_uint64),
        ("nulls", ctypes.c_uint64),
        ("bad", ctypes.c_uint64),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        u32, u64, i64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64
        P = ctypes.POINTER
        L.rm_size_by_sbid.argtypes = [u32, ctypes.c_int]
        L.rm_size_by_sbid.restype = u32
        L.rm_slots_by_size.argtypes = [u32, u32]
        L.rm_slots_by_size.restype = ctypes.c_int
        L.rm_sbid_by_size.argtypes = [u32, u64]
        L.rm_sbid_by_size.restype = ctypes.c_int
        L.rm_header_bytes.argtypes = [u32]
        L.rm_header_bytes.restype = u32
        L.rm_heap_bytes.argtypes = [u32, u32]
        L.rm_heap_bytes.restype = u64
        L.rm_init.argtypes = [P(RmHeap), u32, u32, u64, ctypes.c_void_p]
        L.rm_init.restype = ctypes.c_int
        L.rm_release.argtypes = [P(RmHeap)]
        L.rm_malloc.argtypes = [P(RmHeap), u64, P(ctypes.c_int)]
        L.rm_malloc.restype = u64
        L.rm_free.argtypes = [P(RmHeap), u64]
        L.rm_queue_walk.argtypes = [P(RmHeap)]
        L.rm_queue_walk.restype = i64
        L.rm_replay.argtypes = [P(RmHeap), ctypes.c_void_p, u64, ctypes.c_void_p, ctypes.c_void_p]
        L.rm_replay.restype = i64
        L.rm_mix.argtypes = [u64, u64, u64, u64]
        L.rm_mix.restype = u64
        L.rmt_run.argtypes = [u32, u32, u32, u64, u32, u32, ctypes.c_int, u32, u64, P(RmtResult)]
        L.rmt_run.restype = ctypes.c_int
        _lib = L
    return _lib


def size_by_sbid(log2, sbid):
    return lib().rm_size_by_sbid(log2, sbid)


def slots_by_size(log2, size):
    return lib().rm_slots_by_size(log2, size)


def sbid_by_size(log2, size):
    return lib().rm_sbid_by_size(log2, size)


def header_bytes(log2):
    return lib().rm_header_bytes(log2)


def mix(seed, gid, rnd, k):
    return lib().rm_mix(seed, gid, rnd, k)


class RefHeap:
    """Serialized reference heap (one work-item at a time), exact image."""

    def __init__(self, log2, sblocks, base=0x100000000):
        self.h = RmHeap()
        rc = lib().rm_init(ctypes.byref(self.h), log2, sblocks, base, None)
        if rc != RM_OK:
            raise ValueError("rm_init rejected log2=%d sblocks=%d" % (log2, sblocks))
        self.base = base
        self.log2 = log2
        self.sblocks = sblocks

    def __del__(self):
        try:
            lib().rm_release(ctypes.byref(self.h))
        except Exception:
            pass

    @property
    def nbytes(self):
        return self.h.bytes

    @property
    def hdr(self):
        return self.h.hdr

    def image(self):
        return np.ctypeslib.as_array(self.h.img, shape=(self.h.bytes,)).copy()

    def malloc(self, size):
        st = ctypes.c_int(0)
        p = lib().rm_malloc(ctypes.byref(self.h), size, ctypes.byref(st))
        if st.value == RM_HANG:
            raise RuntimeError("reference would spin forever on malloc(%d)" % size)
        return p

    def free(self, addr):
        lib().rm_free(ctypes.byref(self.h), addr)

    def queue_walk(self):
        return lib().rm_queue_walk(ctypes.byref(self.h))

    def replay(self, ops, pat=None):
        ops = np.ascontiguousarray(ops, dtype=np.int64)
        out = np.zeros(len(ops), dtype=np.uint64)
        patp = None
        if pat is not None:
            pat = np.ascontiguousarray(pat, dtype=np.uint32)
            patp = pat.ctypes.data
        n = lib().rm_replay(ctypes.byref(self.h), ops.ctypes.data, len(ops), out.ctypes.data, patp)
        if n < 0:
            raise RuntimeError("reference would spin forever on this trace")
        return out, n


def cpu_run(log2, sblocks, nthreads, nwi, rounds, mode, fixed=64, seed=1, round0=0):
    r = RmtResult()
    rc = lib().rmt_run(log2, sblocks, nthreads, nwi, rounds, round0, mode, fixed, seed, ctypes.byref(r))
    if rc != RM_OK:
        raise RuntimeError("rmt_run failed")
    return {"seconds": r.seconds, "ops": r.ops, "nulls": r.nulls, "bad": r.bad}
