"""Results do not depend on the tuning: bin thresholds, lanes per vertex,
hub tiles, SSSP near-far bucket width, deferred dynamic order, and the
autotuner's choice."""
import sys

import numpy as np
import pytest

import onegraph_amd as og
import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


TUNINGS = [dict(t1=1, lpv0=1, t2=1, tile=64, delta=0, defer=0, dyn_grid=1),
           dict(t1=4, lpv0=4, t2=64, tile=128, delta=1, defer=1, dyn_grid=0),
           dict(t1=64, lpv0=64, t2=65536, tile=4096, delta=7, defer=0, dyn_grid=3),
           dict(t1=16, lpv0=16, t2=512, tile=512, delta=300, defer=1, dyn_grid=8192),
           dict(t1=8, lpv0=2, t2=1024, tile=256, delta=3, defer=2, dyn_grid=512),
           dict(t1=32, lpv0=8, t2=4096, tile=2048, delta=0, defer=2, dyn_grid=0)]


@pytest.mark.parametrize("approach,frac,k", [("ascetic", 1.0, 1.0), ("ascetic", 0.25, 0.8), ("subway", 0.1, 0.0),
                                              ("ascetic+zerocopy", 0.25, 0.8), ("ascetic+zerocopy+nopack", 0.25, 0.8)])
def test_results_independent_of_tuning(gpu, approach, frac, k):
    """Every tuning gives the oracle's results; the zero-copy plans run the
    fused round relax with every bin width, tile (LDS size) and grid."""
    g = og.Graph.rmat(12, 16, seed=61, weight_seed=62)
    ptr, to, w, _ = g.export()
    E = len(to)
    name, *opts = approach.split("+")
    flags = 0
    for o in opts:
        flags |= {"zerocopy": og.PLAN_ZERO_COPY, "nopack": og.PLAN_NO_PACK}[o]
    g.prepare(name, edge_budget_bytes=int(E * 8 * frac), advisek=k, flags=flags)
    src = int(np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), E)) > 0)[1])
    ref_b, _, _ = oracle.bfs(ptr, to, src)
    ref_s = oracle.sssp(ptr, to, w, src)
    ref_c, _ = oracle.cc(ptr, to)
    for t in TUNINGS:
        _say("tuning", approach, frac, t)
        g.set_tuning(**t)
        assert g.tuning() == t
        assert np.array_equal(g.bfs(src)[0], ref_b)
        d, st = g.sssp(src)
        assert np.array_equal(d, ref_s), t
        assert np.array_equal(g.cc()[0], ref_c)
        lv, cs, _, _ = g.msbfs([src, 0, 5])
        assert np.array_equal(lv[0], ref_b)
    with pytest.raises(og.OgError):
        g.set_tuning(lpv0=3)
    with pytest.raises(og.OgError):
        g.set_tuning(t1=10, t2=5)
    with pytest.raises(og.OgError):
        g.set_tuning(defer=3)
    with pytest.raises(og.OgError):
        g.set_tuning(dyn_grid=70000)
    g.close()


def test_autotune(gpu):
    g = og.Graph.rmat(13, 16, seed=63, weight_seed=64)
    ptr, to, w, _ = g.export()
    g.prepare("ascetic", edge_budget_bytes=len(to) * 8 // 4, advisek=0.8)
    src = int(np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), len(to))) > 0)[0])
    best, ms = g.autotune(og.TUNE_SSSP, src, runs=1)
    assert ms > 0 and g.tuning() == best
    assert np.array_equal(g.sssp(src)[0], oracle.sssp(ptr, to, w, src))
    best, ms = g.autotune(og.TUNE_BFS, src, runs=1)
    assert np.array_equal(g.bfs(src)[0], oracle.bfs(ptr, to, src)[0])
    g.close()


def test_autotune_multi_source(gpu):
    """OG_TUNE_MSBFS tunes on a seeded 64-source batch; the installed tuning
    leaves every source's levels equal to the oracle's (zero-copy plan, so the
    dynamic grid is among the candidates)."""
    g = og.Graph.rmat(13, 16, seed=66)
    ptr, to, _, _ = g.export()
    g.prepare("ascetic", edge_budget_bytes=len(to) * 4 // 2, advisek=0.8, flags=og.PLAN_ZERO_COPY)
    best, ms = g.autotune(og.TUNE_MSBFS, 3004, runs=1)
    assert ms > 0 and g.tuning() == best
    cand = np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), len(to))) > 0)
    srcs = np.random.default_rng(67).choice(cand, 70, replace=False).astype(np.uint32)
    lv, cs, _, _ = g.msbfs(srcs)
    for i in (0, 33, 69):
        ref, _, _ = oracle.bfs(ptr, to, int(srcs[i]))
        assert np.array_equal(lv[i], ref)
        assert og.levels_checksum(ref) == int(cs[i])
    with pytest.raises(og.OgError):
        g.autotune(7, 0)
    g.close()


def test_pagerank_after_retune(gpu):
    g = og.Graph.rmat(11, 16, seed=65).to_pull()
    ptr, cols, _, od = g.export()
    g.prepare("ascetic", edge_budget_bytes=len(cols), advisek=0.8)
    ref, _, _ = oracle.pagerank(od, ptr, cols)
    r1, _ = g.pagerank()
    g.set_tuning(t1=2, t2=8)
    r2, _ = g.pagerank()
    assert np.abs(r1.astype(np.float64) - ref).max() < 1e-6
    assert np.abs(r2.astype(np.float64) - ref).max() < 1e-6
    g.close()
