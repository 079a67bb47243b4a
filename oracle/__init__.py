"""ctypes access to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library
never loads it.  Arrays use the reference CSR layout (ptr has V entries).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")
_lib = None
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.og_oracle_bfs.argtypes = [C.c_uint64, C.c_uint64, _u64p, _u32p, C.c_uint32, _u32p, _u64p, _u32p]
        L.og_oracle_bfs.restype = None
        L.og_oracle_sssp.argtypes = [C.c_uint64, C.c_uint64, _u64p, _u32p, _u32p, C.c_uint32, _u32p]
        L.og_oracle_sssp.restype = None
        L.og_oracle_cc.argtypes = [C.c_uint64, C.c_uint64, _u64p, _u32p, _u32p]
        L.og_oracle_cc.restype = C.c_uint64
        L.og_oracle_pagerank.argtypes = [C.c_uint64, C.c_uint64, _u32p, _u64p, _u32p, C.c_double, C.c_double,
                                         C.c_uint32, _f64p, _u32p, _f64p]
        L.og_oracle_pagerank.restype = None
        L.og_oracle_pagerank_par.argtypes = L.og_oracle_pagerank.argtypes + [C.c_int]
        L.og_oracle_pagerank_par.restype = None
        L.og_oracle_loop_seconds.argtypes = []
        L.og_oracle_loop_seconds.restype = C.c_double
        _lib = L
    return _lib


def _p64(a):
    return a.ctypes.data_as(_u64p)


def _p32(a):
    return a.ctypes.data_as(_u32p)


def _prep(ptr, cols):
    ptr = np.ascontiguousarray(ptr, dtype=np.uint64)
    cols = np.ascontiguousarray(cols, dtype=np.uint32)
    if len(cols) == 0:
        cols = np.zeros(1, np.uint32)[:0]
    return ptr, cols


def bfs(ptr, cols, src):
    """Returns (levels, node_sum, iterations); levels: src = 1, unreached UINT_MAX."""
    ptr, cols = _prep(ptr, cols)
    V, E = len(ptr), len(cols)
    out = np.empty(V, np.uint32)
    ns = C.c_uint64()
    it = C.c_uint32()
    lib().og_oracle_bfs(V, E, _p64(ptr), _p32(np.ascontiguousarray(cols) if E else np.zeros(1, np.uint32)),
                        src, _p32(out), C.byref(ns), C.byref(it))
    return out, ns.value, it.value


def sssp(ptr, to, w, src):
    """Dijkstra; returns distances + 1 (unreached UINT_MAX)."""
    ptr, to = _prep(ptr, to)
    w = np.ascontiguousarray(w, dtype=np.uint32)
    V, E = len(ptr), len(to)
    out = np.empty(V, np.uint32)
    z = np.zeros(1, np.uint32)
    lib().og_oracle_sssp(V, E, _p64(ptr), _p32(to if E else z), _p32(w if E else z), src, _p32(out))
    return out


def cc(ptr, cols):
    """Serial min-label CC; returns (labels, update_count)."""
    ptr, cols = _prep(ptr, cols)
    V, E = len(ptr), len(cols)
    out = np.empty(V, np.uint32)
    n = lib().og_oracle_cc(V, E, _p64(ptr), _p32(cols if E else np.zeros(1, np.uint32)), _p32(out))
    return out, n


def loop_seconds():
    """Wall seconds of the last oracle call's compute loop (allocation and
    initialisation excluded)."""
    return lib().og_oracle_loop_seconds()


def pagerank(out_degree, ptr, in_cols, damping=0.85, tol=1e-7, max_it=100, threads=1, parallel=None):
    """FP64 normalised PageRank over the pull layout; returns (ranks, iterations, l1).
    threads > 1 (or parallel=True) runs og_oracle_pagerank_par, which spreads
    the vertex blocks over threads with the same summation order (identical
    results)."""
    ptr, in_cols = _prep(ptr, in_cols)
    od = np.ascontiguousarray(out_degree, dtype=np.uint32)
    V, E = len(ptr), len(in_cols)
    out = np.empty(V, np.float64)
    it = C.c_uint32()
    l1 = C.c_double()
    args = (V, E, _p32(od), _p64(ptr), _p32(in_cols if E else np.zeros(1, np.uint32)),
            damping, tol, max_it, out.ctypes.data_as(_f64p), C.byref(it), C.byref(l1))
    if parallel or (parallel is None and threads > 1):
        lib().og_oracle_pagerank_par(*args, threads)
    else:
        lib().og_oracle_pagerank(*args)
    return out, it.value, l1.value
