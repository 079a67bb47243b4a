"""Python binding of libonegraph's C ABI (include/onegraph.h) over ctypes.

The engine itself is C++/HIP; this module only marshals numpy arrays and the
ABI structs so tests and bench.py can drive it.  Loading fails loudly when
the shared library has not been built: there is no Python fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libonegraph.so")
CLI_PATH = os.path.join(_HERE, "onegraph")

KIND_PLAIN, KIND_WEIGHTED, KIND_PULL = 0, 1, 2
ASCETIC, SUBWAY, USM = 0, 1, 2
APPROACHES = {"ascetic": ASCETIC, "subway": SUBWAY, "usm": USM}
UNREACHED = 0xFFFFFFFF
PLAN_NO_OVERLAP = 1
PLAN_ZERO_COPY = 2
PLAN_NO_PACK = 4
PLAN_NO_GRAPH = 8
PLAN_DIRECTION_OPT = 16
PLAN_NO_FUSE = 32
EDGES_SYMMETRIZE, EDGES_KEEP_SELF_LOOPS, EDGES_ONE_BASED, EDGES_WEIGHTED = 1, 2, 4, 8

STATUS = {
    0: "OG_OK", 1: "OG_EINVAL", 2: "OG_EIO", 3: "OG_EFORMAT", 4: "OG_EBUDGET",
    5: "OG_ENOMEM_HOST", 6: "OG_ENOMEM_DEVICE", 7: "OG_EHIP", 8: "OG_EBUSY",
    9: "OG_EUNSUPPORTED", 10: "OG_ESTATE", 11: "OG_ETIMEDOUT",
}


class OgError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        self.name = STATUS.get(status, str(status))
        super().__init__(f"{self.name}: {message}")


class PlanParams(C.Structure):
    _fields_ = [("edge_budget_bytes", C.c_uint64), ("advisek", C.c_float), ("param_size", C.c_uint32),
                ("host_threads", C.c_uint32), ("flags", C.c_uint32), ("split_weight", C.c_uint32)]


class PlanInfo(C.Structure):
    _fields_ = [("total_budget_edges", C.c_uint64), ("static_edges", C.c_uint64), ("staging_edges", C.c_uint64),
                ("overload_edges", C.c_uint64), ("static_vertices", C.c_uint64), ("vsplit", C.c_uint64),
                ("device_bytes", C.c_uint64), ("advisek_used", C.c_float), ("approach", C.c_uint32)]


class RunStats(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("approach", C.c_uint32), ("reached", C.c_uint64),
                ("traversed_edges", C.c_uint64), ("work_edges", C.c_uint64), ("updates", C.c_uint64),
                ("node_sum", C.c_uint64), ("static_node_sum", C.c_uint64), ("static_work_edges", C.c_uint64),
                ("h2d_bytes", C.c_uint64), ("chunks", C.c_uint64), ("total_ms", C.c_double),
                ("pre_ms", C.c_double), ("static_ms", C.c_double), ("overload_ms", C.c_double),
                ("overload_move_ms", C.c_double), ("gather_ms", C.c_double),
                ("relax_static_kernel_ms", C.c_double), ("relax_static_launches", C.c_uint64),
                ("relax_static_bytes", C.c_uint64), ("relax_dynamic_kernel_ms", C.c_double),
                ("relax_dynamic_launches", C.c_uint64), ("relax_dynamic_bytes", C.c_uint64),
                ("frontier_kernel_ms", C.c_double), ("frontier_launches", C.c_uint64),
                ("frontier_bytes", C.c_uint64), ("copy_ms", C.c_double), ("checksum", C.c_uint64)]


class GraphInfo(C.Structure):
    _fields_ = [("num_vertices", C.c_uint64), ("num_edges", C.c_uint64), ("kind", C.c_uint32),
                ("max_degree", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("t1", C.c_uint32), ("lpv0", C.c_uint32), ("t2", C.c_uint32), ("tile", C.c_uint32),
                ("delta", C.c_uint32), ("defer", C.c_uint32), ("dyn_grid", C.c_uint32)]


TUNE_BFS, TUNE_SSSP, TUNE_CC, TUNE_MSBFS = 0, 1, 2, 3


class PrParams(C.Structure):
    _fields_ = [("damping", C.c_double), ("tolerance", C.c_double), ("max_iterations", C.c_uint32),
                ("fp64", C.c_uint32)]


class PrStats(C.Structure):
    _fields_ = [("run", RunStats), ("iterations", C.c_uint32), ("final_l1", C.c_double)]


SERVER_MAX_GRAPHS = 8
QUERY_BFS, QUERY_SSSP = 0, 1
QUERY_KEEP_VALUES = 1
PATH_BATCH, PATH_LANE = 0, 1  # og_query_result.path: a batch engine's slot / a single-source SSSP lane
SERVER_NO_SSSP_LANES = 0x1  # every SSSP query runs in the batched engine (og.h OG_SERVER_NO_SSSP_LANES)
SERVER_SSSP_BATCH = 0x2  # the batched SSSP engine beside the lanes (og.h OG_SERVER_SSSP_BATCH)
SERVER_TEST_FAIL_SSSP = 0x80000000  # testing: SSSP engines stop at their first round (og.h OG_SERVER_TEST_FAIL_SSSP)


class ServerParams(C.Structure):
    _fields_ = [("state_budget_bytes", C.c_uint64), ("max_pending", C.c_uint32), ("flags", C.c_uint32),
                ("deadline_ms", C.c_uint32), ("sssp_delta", C.c_uint32), ("sssp_lanes", C.c_uint32),
                ("sssp_running_max", C.c_uint32)]


class QueryResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("graph", C.c_uint32), ("kind", C.c_uint32), ("source", C.c_uint32),
                ("status", C.c_uint32), ("reached", C.c_uint64), ("traversed_edges", C.c_uint64),
                ("checksum", C.c_uint64), ("queue_ms", C.c_double), ("latency_ms", C.c_double),
                ("rounds", C.c_uint32), ("path", C.c_uint32)]


class ServerInfo(C.Structure):
    _fields_ = [("num_graphs", C.c_uint32), ("bfs_slots", C.c_uint32 * SERVER_MAX_GRAPHS),
                ("sssp_slots", C.c_uint32 * SERVER_MAX_GRAPHS), ("state_bytes", C.c_uint64),
                ("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("pending", C.c_uint64), ("in_flight", C.c_uint64), ("rounds", C.c_uint64)]


def as_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        if isinstance(v, C.Structure):
            v = as_dict(v)
        elif isinstance(v, C.Array):
            v = list(v)
        out[name] = v
    return out


_lib = None
_P = C.c_void_p
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


def lib():
    """Loads libonegraph.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m onegraph_amd.build`")
    L = C.CDLL(LIB_PATH)
    sig = {
        "og_last_error": (C.c_char_p, []),
        "og_version": (C.c_char_p, []),
        "og_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "og_graph_load": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_P)]),
        "og_graph_save": (C.c_int, [_P, C.c_char_p]),
        "og_graph_from_csr": (C.c_int, [C.c_uint64, C.c_uint64, _u64p, _u32p, _u32p, _u32p, C.c_int, C.POINTER(_P)]),
        "og_graph_from_edges": (C.c_int, [C.c_uint64, C.c_uint64, _u32p, _u32p, _u32p, C.c_uint32, C.POINTER(_P)]),
        "og_graph_load_edgelist": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(_P)]),
        "og_graph_generate_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                             C.c_uint64, C.c_uint64, C.POINTER(_P)]),
        "og_graph_generate_er": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint64, C.POINTER(_P)]),
        "og_graph_to_pull": (C.c_int, [_P, C.POINTER(_P)]),
        "og_graph_add_weights": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
        "og_graph_get_info": (C.c_int, [_P, C.POINTER(GraphInfo)]),
        "og_graph_export": (C.c_int, [_P, _u64p, _u32p, _u32p, _u32p]),
        "og_graph_find_big_node": (C.c_int, [_P, C.c_uint32, _u32p]),
        "og_graph_free": (C.c_int, [_P]),
        "og_plan_compute": (C.c_int, [C.c_uint64, C.c_uint64, _u64p, C.c_uint32, C.POINTER(PlanParams),
                                      C.POINTER(PlanInfo)]),
        "og_prepare": (C.c_int, [_P, C.c_int, C.POINTER(PlanParams), C.POINTER(PlanInfo)]),
        "og_release": (C.c_int, [_P]),
        "og_bfs": (C.c_int, [_P, C.c_uint32, _u32p, C.POINTER(RunStats)]),
        "og_sssp": (C.c_int, [_P, C.c_uint32, _u32p, C.POINTER(RunStats)]),
        "og_cc": (C.c_int, [_P, _u32p, C.POINTER(RunStats)]),
        "og_msbfs": (C.c_int, [_P, _u32p, C.c_uint32, _u32p, _u64p, _u64p, C.POINTER(RunStats)]),
        "og_levels_checksum": (C.c_int, [_u32p, C.c_uint64, _u64p]),
        "og_bfs_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_void_p,
                                            C.c_void_p, C.POINTER(RunStats)]),
        "og_probe_hbm_gather": (C.c_int, [C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "og_probe_host_link": (C.c_int, [C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]),
        "og_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
        "og_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
        "og_autotune": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(Tuning), C.POINTER(C.c_double)]),
        "og_server_create": (C.c_int, [C.POINTER(_P), C.c_uint32, C.POINTER(ServerParams), C.POINTER(_P)]),
        "og_server_submit": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _u64p]),
        "og_server_poll": (C.c_int, [_P, C.POINTER(QueryResult), C.c_uint32, _u32p, C.c_int32]),
        "og_server_take_values": (C.c_int, [_P, C.c_uint64, _u32p]),
        "og_server_get_info": (C.c_int, [_P, C.POINTER(ServerInfo)]),
        "og_server_destroy": (C.c_int, [_P]),
        "og_pagerank": (C.c_int, [_P, C.POINTER(PrParams), C.POINTER(C.c_float), C.POINTER(C.c_double),
                                  C.POINTER(PrStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(st):
    if st != 0:
        raise OgError(st, lib().og_last_error().decode(errors="replace"))


def _u32(a):
    return a.ctypes.data_as(_u32p) if a is not None else None


def _u64(a):
    return a.ctypes.data_as(_u64p) if a is not None else None


def device_count():
    n = C.c_int(0)
    _check(lib().og_device_count(C.byref(n)))
    return n.value


def bfs_device_noalloc(row_ptr, cols, num_vertices, num_edges, source, levels, stream=None):
    """og_bfs_device_noalloc on device addresses (ints): row_ptr holds V + 1
    u64 offsets, cols E u32 ids, levels receives V u32 levels.  Returns the
    run statistics."""
    st = RunStats()
    _check(lib().og_bfs_device_noalloc(C.c_void_p(row_ptr), C.c_void_p(cols), num_vertices, num_edges, source,
                                       C.c_void_p(levels), C.c_void_p(stream), C.byref(st)))
    return as_dict(st)


def levels_checksum(values):
    """Digest of a BFS/SSSP result array (same as og_run_stats.checksum)."""
    v = np.ascontiguousarray(values, dtype=np.uint32)
    out = C.c_uint64()
    _check(lib().og_levels_checksum(_u32(v), len(v), C.byref(out)))
    return out.value


def probe_host_link(nbytes=1 << 30):
    """GB/s of shader reads of pinned host memory (contiguous and scattered
    short rows) and of hipMemcpy host-to-device, on the current device."""
    c, sc, mc = C.c_double(), C.c_double(), C.c_double()
    _check(lib().og_probe_host_link(nbytes, C.byref(c), C.byref(sc), C.byref(mc)))
    return {"contiguous_gbs": c.value, "scattered_rows_gbs": sc.value, "memcpy_gbs": mc.value}


def probe_hbm_gather(table_bytes):
    """Random 4-byte reads from a device table of table_bytes: useful GB/s and
    million reads per second (the random-access ceiling of the relax)."""
    g, m = C.c_double(), C.c_double()
    _check(lib().og_probe_hbm_gather(table_bytes, C.byref(g), C.byref(m)))
    return {"table_bytes": table_bytes, "gather_gbs": g.value, "mgathers_per_s": m.value}


def version():
    return lib().og_version().decode()


def plan_compute(ptr, num_edges, edge_bytes, edge_budget_bytes=0, advisek=0.0, param_size=0):
    """Static-region plan for a CSR with V-entry offsets (reference layout)."""
    ptr = np.ascontiguousarray(ptr, dtype=np.uint64)
    p = PlanParams(edge_budget_bytes, advisek, param_size, 0, 0)
    info = PlanInfo()
    _check(lib().og_plan_compute(len(ptr), num_edges, _u64(ptr), edge_bytes, C.byref(p), C.byref(info)))
    return as_dict(info)


class Graph:
    """Owns an og_graph handle (host CSR in pinned memory + device state)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    # ------------------------------------------------------------ creation
    @classmethod
    def _make(cls, fn, *args):
        h = _P()
        _check(fn(*args, C.byref(h)))
        return cls(h.value)

    @classmethod
    def load(cls, path, kind=KIND_PLAIN):
        return cls._make(lib().og_graph_load, path.encode(), kind)

    @classmethod
    def from_csr(cls, ptr, cols, weights=None, out_degree=None, kind=None):
        """ptr has V entries (reference layout: no sentinel, row V-1 ends at E)."""
        ptr = np.ascontiguousarray(ptr, dtype=np.uint64)
        cols = np.ascontiguousarray(cols, dtype=np.uint32)
        E = len(cols)
        if kind is None:
            kind = KIND_PULL if out_degree is not None else (KIND_WEIGHTED if weights is not None else KIND_PLAIN)
        w = np.ascontiguousarray(weights, dtype=np.uint32) if weights is not None else None
        od = np.ascontiguousarray(out_degree, dtype=np.uint32) if out_degree is not None else None
        V = len(ptr)
        return cls._make(lib().og_graph_from_csr, V, E, _u64(ptr), _u32(cols) if E else _u32(np.zeros(1, np.uint32)),
                         _u32(w), _u32(od), kind)

    @classmethod
    def from_edges(cls, src, dst, weights=None, num_vertices=0, symmetrize=False, keep_self_loops=False):
        """Arcs src[i] -> dst[i]: deduplicated (a repeated arc keeps its
        smallest weight), self-loops dropped unless kept, rows sorted."""
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        if len(src) != len(dst):
            raise ValueError("src and dst differ in length")
        w = np.ascontiguousarray(weights, dtype=np.uint32) if weights is not None else None
        if w is not None and len(w) != len(src):
            raise ValueError("weights differ in length")
        flags = (EDGES_SYMMETRIZE if symmetrize else 0) | (EDGES_KEEP_SELF_LOOPS if keep_self_loops else 0)
        z = np.zeros(1, np.uint32)
        return cls._make(lib().og_graph_from_edges, num_vertices, len(src), _u32(src if len(src) else z),
                         _u32(dst if len(dst) else z), _u32(w) if w is None or len(w) else _u32(z), flags)

    @classmethod
    def load_edgelist(cls, path, weighted=False, symmetrize=False, keep_self_loops=False, one_based=False):
        """Text edge list ("u v [w]" lines, '#'/'%' comments) or MatrixMarket."""
        flags = ((EDGES_WEIGHTED if weighted else 0) | (EDGES_SYMMETRIZE if symmetrize else 0) |
                 (EDGES_KEEP_SELF_LOOPS if keep_self_loops else 0) | (EDGES_ONE_BASED if one_based else 0))
        return cls._make(lib().og_graph_load_edgelist, str(path).encode(), flags)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, weight_seed=0):
        return cls._make(lib().og_graph_generate_rmat, scale, edge_factor, a, b, c, seed, weight_seed)

    @classmethod
    def er(cls, n, avg_degree, seed=1):
        return cls._make(lib().og_graph_generate_er, n, avg_degree, seed)

    def to_pull(self):
        return Graph._make(lib().og_graph_to_pull, self._h)

    def add_weights(self, seed):
        return Graph._make(lib().og_graph_add_weights, self._h, seed)

    # ------------------------------------------------------------ queries
    def info(self):
        i = GraphInfo()
        _check(lib().og_graph_get_info(self._h, C.byref(i)))
        return as_dict(i)

    @property
    def num_vertices(self):
        return self.info()["num_vertices"]

    @property
    def num_edges(self):
        return self.info()["num_edges"]

    def export(self):
        """Returns (ptr[V], cols[E], weights or None, out_degree or None)."""
        inf = self.info()
        V, E, kind = inf["num_vertices"], inf["num_edges"], inf["kind"]
        ptr = np.empty(V, np.uint64)
        cols = np.empty(max(E, 1), np.uint32)
        w = np.empty(max(E, 1), np.uint32) if kind == KIND_WEIGHTED else None
        od = np.empty(V, np.uint32) if kind == KIND_PULL else None
        _check(lib().og_graph_export(self._h, _u64(ptr), _u32(cols), _u32(w), _u32(od)))
        return ptr, cols[:E], (w[:E] if w is not None else None), od

    def save(self, path):
        _check(lib().og_graph_save(self._h, path.encode()))

    def find_big_node(self, last=0):
        out = C.c_uint32()
        _check(lib().og_graph_find_big_node(self._h, last, C.byref(out)))
        return out.value

    # ------------------------------------------------------------ device
    def prepare(self, approach=ASCETIC, edge_budget_bytes=0, advisek=0.0, param_size=0, host_threads=0, flags=0,
                split_weight=0):
        if isinstance(approach, str):
            approach = APPROACHES[approach]
        p = PlanParams(edge_budget_bytes, advisek, param_size, host_threads, flags, split_weight)
        info = PlanInfo()
        _check(lib().og_prepare(self._h, approach, C.byref(p), C.byref(info)))
        return as_dict(info)

    def release(self):
        _check(lib().og_release(self._h))

    def tuning(self):
        t = Tuning()
        _check(lib().og_get_tuning(self._h, C.byref(t)))
        return as_dict(t)

    def set_tuning(self, **kw):
        t = Tuning(**{**self.tuning(), **kw})
        _check(lib().og_set_tuning(self._h, C.byref(t)))
        return as_dict(t)

    def autotune(self, algo=TUNE_BFS, source=0, runs=3):
        t = Tuning()
        ms = C.c_double()
        _check(lib().og_autotune(self._h, algo, source, runs, C.byref(t), C.byref(ms)))
        return as_dict(t), ms.value

    def bfs(self, source, want_values=True):
        out = np.empty(self.num_vertices, np.uint32) if want_values else None
        st = RunStats()
        _check(lib().og_bfs(self._h, source, _u32(out), C.byref(st)))
        return out, as_dict(st)

    def sssp(self, source, want_values=True):
        out = np.empty(self.num_vertices, np.uint32) if want_values else None
        st = RunStats()
        _check(lib().og_sssp(self._h, source, _u32(out), C.byref(st)))
        return out, as_dict(st)

    def msbfs(self, sources, want_levels=True):
        """Multi-source BFS; returns (levels [n, V] or None, checksums, reached, stats)."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        n, V = len(src), self.num_vertices
        lv = np.empty((n, V), np.uint32) if want_levels else None
        cs = np.empty(n, np.uint64)
        rc = np.empty(n, np.uint64)
        st = RunStats()
        _check(lib().og_msbfs(self._h, _u32(src), n, _u32(lv), _u64(cs), _u64(rc), C.byref(st)))
        return lv, cs, rc, as_dict(st)

    def cc(self, want_values=True):
        out = np.empty(self.num_vertices, np.uint32) if want_values else None
        st = RunStats()
        _check(lib().og_cc(self._h, _u32(out), C.byref(st)))
        return out, as_dict(st)

    def pagerank(self, damping=0.85, tolerance=1e-7, max_iterations=100, fp64=False, want_values=True):
        V = self.num_vertices
        f32 = np.empty(V, np.float32) if (want_values and not fp64) else None
        f64 = np.empty(V, np.float64) if (want_values and fp64) else None
        p = PrParams(damping, tolerance, max_iterations, 1 if fp64 else 0)
        st = PrStats()
        _check(lib().og_pagerank(self._h, C.byref(p),
                                 f32.ctypes.data_as(C.POINTER(C.c_float)) if f32 is not None else None,
                                 f64.ctypes.data_as(C.POINTER(C.c_double)) if f64 is not None else None,
                                 C.byref(st)))
        return (f64 if fp64 else f32), as_dict(st)

    # ------------------------------------------------------------ lifetime
    def close(self):
        if self._h and self._h.value:
            _check(lib().og_graph_free(self._h))
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Continuous-batching query server over device-resident graphs."""

    def __init__(self, graphs, state_budget_bytes=0, max_pending=0, flags=0, sssp_delta=0, deadline_ms=0,
                 sssp_lanes=0, sssp_running_max=0):
        self._graphs = list(graphs)
        arr = (_P * len(self._graphs))(*[g._h.value for g in self._graphs])
        p = ServerParams(state_budget_bytes, max_pending, flags, deadline_ms, sssp_delta, sssp_lanes, sssp_running_max)
        h = _P()
        _check(lib().og_server_create(arr, len(self._graphs), C.byref(p), C.byref(h)))
        self._h = h
        self._V = [g.num_vertices for g in self._graphs]

    def submit(self, graph, kind, source, keep=False):
        qid = C.c_uint64()
        _check(lib().og_server_submit(self._h, graph, kind, source, QUERY_KEEP_VALUES if keep else 0,
                                      C.byref(qid)))
        return qid.value

    def poll(self, max_results=1024, timeout_ms=0):
        buf = (QueryResult * max_results)()
        n = C.c_uint32()
        _check(lib().og_server_poll(self._h, buf, max_results, C.byref(n), timeout_ms))
        return [as_dict(buf[i]) for i in range(n.value)]

    def take_values(self, qid, graph):
        out = np.empty(self._V[graph], np.uint32)
        _check(lib().og_server_take_values(self._h, qid, _u32(out)))
        return out

    def info(self):
        i = ServerInfo()
        _check(lib().og_server_get_info(self._h, C.byref(i)))
        d = as_dict(i)
        d["bfs_slots"] = d["bfs_slots"][: d["num_graphs"]]
        d["sssp_slots"] = d["sssp_slots"][: d["num_graphs"]]
        return d

    def close(self):
        if self._h and self._h.value:
            _check(lib().og_server_destroy(self._h))
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
