"""og_bfs_device_noalloc, the allocation-free variant of og_bfs: on CSRs
placed in device memory by the test, its levels and statistics must equal
og_bfs's on the same graph and source (and the oracle's), on consecutive
calls whose second input is smaller, since every call shares one static
scratch array on the GPU."""
import ctypes as C

import numpy as np
import pytest

import onegraph_amd as og
import oracle

pytestmark = pytest.mark.gpu


class _Hip:
    """The HIP runtime libonegraph.so links (same loaded instance)."""

    def __init__(self):
        og.lib()
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        self.L = L
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 4)) == 0
        self.bufs.append(p)
        return p.value

    def download(self, p, dtype, n):
        out = np.empty(n, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, C.c_void_p(p), out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.L.hipDeviceSynchronize()
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _case(hip, scale, ef, seed, pick):
    g = og.Graph.rmat(scale, ef, seed=seed)
    ptr, cols, _, _ = g.export()
    V, E = len(ptr), len(cols)
    deg = np.diff(np.append(ptr.astype(np.int64), E))
    src = int(np.flatnonzero(deg > 0)[pick]) if pick >= 0 else int(np.flatnonzero(deg == 0)[0])
    d_ptr = hip.upload(np.append(ptr, np.uint64(E)).astype(np.uint64))
    d_cols = hip.upload(cols)
    d_lv = hip.empty(4 * V)
    print(f"[case] noalloc rmat{scale} ef{ef} V={V} E={E} src={src}", flush=True)
    st = og.bfs_device_noalloc(d_ptr, d_cols, V, E, src, d_lv)
    got = hip.download(d_lv, np.uint32, V)
    g.prepare("ascetic", edge_budget_bytes=E * 4, advisek=1.0)
    ref_lv, ref_st = g.bfs(src)
    g.close()
    return got, st, ref_lv, ref_st, ptr, cols, src


def test_noalloc_equals_og_bfs_twice_then_smaller(gpu):
    hip = _Hip()
    try:
        # scale 16 has hub rows (degree > 4096: LDS hub tiles); then a
        # smaller graph, then a source with no out-edges
        for scale, ef, seed, pick in ((16, 16, 41, 0), (12, 8, 42, 3), (11, 4, 43, -1)):
            got, st, ref_lv, ref_st, ptr, cols, src = _case(hip, scale, ef, seed, pick)
            assert np.array_equal(got, ref_lv), scale
            for k in ("reached", "traversed_edges", "checksum", "iterations", "node_sum", "work_edges"):
                assert st[k] == ref_st[k], (scale, k)
            oref, _, _ = oracle.bfs(ptr, cols, src)
            assert np.array_equal(got, oref)
    finally:
        hip.free()


def test_noalloc_rejects_what_its_scratch_cannot_hold(gpu):
    hip = _Hip()
    try:
        d = hip.empty(64)
        for V, E in ((1 << 21 | 1, 4), (16, (1 << 26) + 1)):
            with pytest.raises(og.OgError) as e:
                og.bfs_device_noalloc(d, d, V, E, 0, d)
            assert e.value.name == "OG_EINVAL" and "scratch" in str(e.value)
        with pytest.raises(og.OgError) as e:  # source out of range
            og.bfs_device_noalloc(d, d, 4, 0, 4, d)
        assert e.value.name == "OG_EINVAL"
        # row offsets that do not end at E (a V-entry array without the sentinel)
        ptr = hip.upload(np.array([0, 1, 2, 3, 3], np.uint64))
        cols = hip.upload(np.array([1, 2, 3], np.uint32))
        lv = hip.empty(16)
        with pytest.raises(og.OgError) as e:
            og.bfs_device_noalloc(ptr, cols, 4, 4, 0, lv)
        assert e.value.name == "OG_EINVAL"
        st = og.bfs_device_noalloc(ptr, cols, 4, 3, 0, lv)  # the consistent call works after the refusals
        assert hip.download(lv, np.uint32, 4).tolist() == [1, 2, 3, 4] and st["reached"] == 4
    finally:
        hip.free()
