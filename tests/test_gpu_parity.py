"""GPU parity: every approach and plan against the CPU oracle (bit-exact for
BFS/SSSP/CC, max-abs 1e-6 for PageRank).  Runs on the MI355X box only."""
import sys

import numpy as np
import pytest

import onegraph_amd as og
import oracle

pytestmark = pytest.mark.gpu

# (approach, edge budget as a fraction of the CSR edge bytes, advisek)
PLANS = [
    ("ascetic", 1.0, 1.0),     # whole graph in HBM
    ("ascetic", 0.25, 0.8),    # the out-of-memory configuration of the benchmarks
    ("ascetic", 1 / 64, 0.5),  # tiny staging: many chunks through both slots
    ("ascetic", 0.25, 0.0),    # reference auto-advisek rule
    ("ascetic+zerocopy", 0.25, 0.8),  # dynamic edges read in place over the link
    ("ascetic+zerocopy", 1 / 64, 0.5),  # small HBM-resident head, most dynamic reads over the link
    ("ascetic+zerocopy+nopack", 1 / 8, 0.5),  # 8-byte rows read in place
    ("ascetic+zerocopy+nograph", 0.25, 0.8),  # device-driven rounds enqueued eagerly
    ("ascetic+zerocopy+nofuse", 0.25, 0.8),  # static and dynamic relax on two streams
    ("ascetic+zerocopy+nofuse", 1 / 64, 0.5),
    ("ascetic+nograph", 1.0, 1.0),
    ("ascetic+nopack", 1 / 64, 0.5),  # 8-byte weighted edges in HBM and on the link
    ("subway+zerocopy", 0.25, 0.0),
    ("subway", 0.25, 0.0),
    ("subway", 1 / 64, 0.0),
    ("usm", 0.0, 0.0),
]


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _prepare(g, approach, frac, k, edge_bytes):
    E = g.num_edges
    budget = int(max(1, E * edge_bytes * frac)) if frac else 0
    name, *opts = approach.split("+")
    flags = 0
    for o in opts:
        flags |= {"zerocopy": og.PLAN_ZERO_COPY, "nopack": og.PLAN_NO_PACK, "nograph": og.PLAN_NO_GRAPH,
                  "nooverlap": og.PLAN_NO_OVERLAP, "diropt": og.PLAN_DIRECTION_OPT, "nofuse": og.PLAN_NO_FUSE}[o]
    return g.prepare(name, edge_budget_bytes=budget, advisek=k, flags=flags)


def _sources(ptr, E, n, seed):
    deg = np.diff(np.append(ptr.astype(np.int64), E))
    cand = np.flatnonzero(deg > 0)
    rng = np.random.default_rng(seed)
    return [int(cand[0])] + [int(x) for x in rng.choice(cand, size=min(n, len(cand)), replace=False)]


@pytest.fixture(scope="module")
def rmat_w():
    g = og.Graph.rmat(13, 16, seed=11, weight_seed=12)
    yield g
    g.close()


@pytest.mark.parametrize("approach,frac,k", PLANS)
def test_bfs_rmat(gpu, rmat_w, approach, frac, k):
    ptr, to, w, _ = rmat_w.export()
    _say("bfs rmat13", approach, frac, k)
    plan = _prepare(rmat_w, approach, frac, k, 8)
    for s in _sources(ptr, len(to), 3, 5):
        lv, st = rmat_w.bfs(s)
        ref, ns, it = oracle.bfs(ptr, to, s)
        assert np.array_equal(lv, ref), (approach, s)
        assert st["reached"] == int((ref != og.UNREACHED).sum())
        assert st["node_sum"] == ns  # level-synchronous: same frontier sizes as runCPUBFS
    if approach == "ascetic" and frac < 1:
        assert plan["vsplit"] < rmat_w.num_vertices


@pytest.mark.parametrize("approach,frac,k", PLANS)
def test_sssp_rmat(gpu, rmat_w, approach, frac, k):
    ptr, to, w, _ = rmat_w.export()
    _say("sssp rmat13", approach, frac, k)
    _prepare(rmat_w, approach, frac, k, 8)
    for s in _sources(ptr, len(to), 3, 6):
        d, st = rmat_w.sssp(s)
        assert np.array_equal(d, oracle.sssp(ptr, to, w, s)), (approach, s)
        if approach == "ascetic" and frac < 1:
            assert st["h2d_bytes"] > 0


@pytest.mark.parametrize("approach,frac,k", PLANS)
def test_cc_rmat(gpu, rmat_w, approach, frac, k):
    ptr, to, w, _ = rmat_w.export()
    _say("cc rmat13", approach, frac, k)
    _prepare(rmat_w, approach, frac, k, 8)
    lab, st = rmat_w.cc()
    ref, _ = oracle.cc(ptr, to)
    assert np.array_equal(lab, ref), approach


@pytest.mark.parametrize("approach,frac,k", [p for p in PLANS if p[1] != 1 / 64 or p[0] == "ascetic"])
def test_cc_er(gpu, approach, frac, k):
    g = og.Graph.er(1 << 12, 8, seed=21)
    ptr, to, _, _ = g.export()
    _say("cc er12", approach, frac, k)
    _prepare(g, approach, frac, k, 4)
    lab, _ = g.cc()
    ref, _ = oracle.cc(ptr, to)
    assert np.array_equal(lab, ref)
    g.close()


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("approach,frac,k", [("ascetic", 1.0, 1.0), ("ascetic", 0.25, 0.8), ("subway", 0.25, 0.0),
                                              ("usm", 0.0, 0.0)])
def test_pagerank(gpu, approach, frac, k, fp64):
    g0 = og.Graph.rmat(12, 16, seed=31)
    g = g0.to_pull()
    ptr, cols, _, od = g.export()
    _say("pagerank rmat12", approach, frac, k, "fp64" if fp64 else "fp32")
    _prepare(g, approach, frac, k, 4)
    r, st = g.pagerank(fp64=fp64)
    ref, it, l1 = oracle.pagerank(od, ptr, cols)
    err = np.abs(r.astype(np.float64) - ref).max()
    assert err < 1e-6, err
    assert abs(st["iterations"] - it) <= 1
    assert abs(float(np.sum(r, dtype=np.float64)) - 1.0) < 1e-5
    g.close()
    g0.close()


def test_small_fixtures(gpu):
    # path 0->1->2->3, plus an isolated vertex 4 and a self-contained pair 5<->6
    ptr = np.array([0, 1, 2, 3, 3, 3, 4], np.uint64)
    cols = np.array([1, 2, 3, 6, 5], np.uint32)
    w = np.array([5, 1, 63, 2, 2], np.uint32)
    g = og.Graph.from_csr(ptr, cols, weights=w)
    for approach, budget in [("ascetic", 32), ("ascetic", 0), ("subway", 16), ("usm", 0)]:
        _say("fixture path", approach, budget)
        g.prepare(approach, edge_budget_bytes=budget, advisek=0.5)
        lv, _ = g.bfs(0)
        assert lv.tolist() == [1, 2, 3, 4, og.UNREACHED, og.UNREACHED, og.UNREACHED]
        d, _ = g.sssp(0)
        assert d.tolist() == [1, 6, 7, 70, og.UNREACHED, og.UNREACHED, og.UNREACHED]
        lab, _ = g.cc()
        assert np.array_equal(lab, oracle.cc(ptr, cols)[0])
    g.close()


def test_errors(gpu):
    g = og.Graph.rmat(8, 4, seed=1)
    with pytest.raises(og.OgError) as e:
        g.bfs(0)
    assert e.value.name == "OG_ESTATE"
    g.prepare("ascetic")
    with pytest.raises(og.OgError) as e:
        g.bfs(1 << 20)
    assert e.value.name == "OG_EINVAL"
    with pytest.raises(og.OgError) as e:
        g.sssp(0)
    assert e.value.name == "OG_EUNSUPPORTED"
    g.close()


@pytest.mark.parametrize("approach,frac,k", [("ascetic", 1.0, 1.0), ("ascetic", 0.25, 0.8), ("ascetic", 1 / 64, 0.5),
                                              ("ascetic+zerocopy", 0.25, 0.8), ("subway", 0.25, 0.0),
                                              ("usm", 0.0, 0.0)])
def test_msbfs(gpu, rmat_w, approach, frac, k):
    ptr, to, w, _ = rmat_w.export()
    _say("msbfs rmat13", approach, frac, k)
    _prepare(rmat_w, approach, frac, k, 8)
    deg = np.diff(np.append(ptr.astype(np.int64), len(to)))
    rng = np.random.default_rng(41)
    srcs = rng.choice(np.flatnonzero(deg > 0), 100, replace=False).astype(np.uint32)
    srcs[5] = srcs[6]  # a repeated source inside one batch
    srcs[70] = 0       # a source with out-degree 0 in the second batch
    lv, cs, rc, st = rmat_w.msbfs(srcs)
    for i, s in enumerate(srcs):
        ref, _, _ = oracle.bfs(ptr, to, int(s))
        assert np.array_equal(lv[i], ref), (approach, i, int(s))
        assert cs[i] == og.levels_checksum(ref)
        assert rc[i] == int((ref != og.UNREACHED).sum())
    # single-source runs agree through the device digest alone
    for s in srcs[:3]:
        _, st1 = rmat_w.bfs(int(s), want_values=False)
        assert st1["checksum"] == cs[list(srcs).index(s)]
    _, cs2, _, _ = rmat_w.msbfs(srcs[:64], want_levels=False)
    assert np.array_equal(cs2, cs[:64])


@pytest.mark.parametrize("approach,frac,k", [("ascetic", 1.0, 1.0), ("ascetic+zerocopy", 0.25, 0.8),
                                              ("ascetic+zerocopy+nooverlap", 0.25, 0.8), ("usm", 0.0, 0.0)])
def test_graph_replays(gpu, rmat_w, approach, frac, k):
    """Device-driven rounds replayed from a captured hipGraph: every replay of
    every call equals the eagerly enqueued rounds and the oracle, with other
    runs (another algorithm, another source) in between."""
    ptr, to, w, _ = rmat_w.export()
    srcs = _sources(ptr, len(to), 2, 9)
    _say("graph replays rmat13", approach, frac, k)
    _prepare(rmat_w, approach + "+nograph", frac, k, 8)
    eager = {s: (rmat_w.bfs(s)[0], rmat_w.sssp(s)[0]) for s in srcs}
    eager_cc = rmat_w.cc()[0]
    _prepare(rmat_w, approach, frac, k, 8)
    for rep in range(3):
        for s in srcs:
            lv, _ = rmat_w.bfs(s)
            d, _ = rmat_w.sssp(s)
            assert np.array_equal(lv, eager[s][0]), (rep, s)
            assert np.array_equal(d, eager[s][1]), (rep, s)
        assert np.array_equal(rmat_w.cc()[0], eager_cc), rep
    for s in srcs:
        assert np.array_equal(eager[s][0], oracle.bfs(ptr, to, s)[0])
        assert np.array_equal(eager[s][1], oracle.sssp(ptr, to, w, s))
    assert np.array_equal(eager_cc, oracle.cc(ptr, to)[0])
    # a retune drops the captured rounds; the next capture must agree too
    rmat_w.set_tuning(t1=16, lpv0=4, delta=8)
    for s in srcs:
        assert np.array_equal(rmat_w.sssp(s)[0], eager[s][1])


@pytest.mark.parametrize("approach,frac,k", [("ascetic+zerocopy", 0.25, 0.8), ("ascetic", 1 / 64, 0.5),
                                              ("ascetic+zerocopy", 1 / 64, 0.5), ("subway+zerocopy", 0.25, 0.0)])
@pytest.mark.parametrize("delta,defer", [(0, 1), (8, 1), (1, 2), (8, 2), (300, 2)])
def test_deferred_dynamic_order(gpu, rmat_w, approach, frac, k, delta, defer):
    """Holding the dynamic list back while static vertices are active (defer
    1), or until the static part of its near-far bucket has settled (defer 2,
    SSSP), changes the relax order, not the fixed point: SSSP and CC stay
    bit-exact."""
    ptr, to, w, _ = rmat_w.export()
    _say("deferred dynamic rmat13", approach, frac, k, delta, defer)
    _prepare(rmat_w, approach, frac, k, 8)
    rmat_w.set_tuning(defer=defer, delta=delta)
    for s in _sources(ptr, len(to), 2, 13):
        d, st = rmat_w.sssp(s)
        assert np.array_equal(d, oracle.sssp(ptr, to, w, s)), s
    assert np.array_equal(rmat_w.cc()[0], oracle.cc(ptr, to)[0])
    lv, _ = rmat_w.bfs(3)  # BFS ignores defer (plain level stores)
    assert np.array_equal(lv, oracle.bfs(ptr, to, 3)[0])


@pytest.mark.parametrize("graph", ["rmat13", "rmat16", "er12", "chain"])
@pytest.mark.parametrize("graphmode", ["", "+nograph"])
def test_direction_optimising_bfs(gpu, graph, graphmode):
    """Bottom-up rounds give the same levels and the same per-round frontier
    sizes (node_sum) as the level-synchronous oracle; other algorithms on the
    same prepared graph are unaffected."""
    if graph == "rmat13":
        g = og.Graph.rmat(13, 16, seed=71, weight_seed=72)
    elif graph == "rmat16":
        g = og.Graph.rmat(16, 16, seed=73)
    elif graph == "er12":
        g = og.Graph.er(1 << 12, 16, seed=74)
    else:  # long directed chain with a shortcut star: many rounds, tiny frontiers
        n = 3000
        src = np.concatenate([np.arange(n - 1), np.zeros(100, np.int64)])
        dst = np.concatenate([np.arange(1, n), np.arange(1000, 1100)])
        g = og.Graph.from_edges(src, dst, num_vertices=n + 5)
    ptr, to, w, _ = g.export()
    _say("direction-optimising bfs", graph, graphmode)
    _prepare(g, "ascetic+diropt" + graphmode, 1.0, 1.0, 8 if w is not None else 4)
    for s in _sources(ptr, len(to), 3, 15):
        lv, st = g.bfs(s)
        ref, ns, it = oracle.bfs(ptr, to, s)
        assert np.array_equal(lv, ref), (graph, s)
        assert st["node_sum"] == ns
        assert st["reached"] == int((ref != og.UNREACHED).sum())
    if w is not None:
        s = _sources(ptr, len(to), 1, 16)[1]
        assert np.array_equal(g.sssp(s)[0], oracle.sssp(ptr, to, w, s))
    assert np.array_equal(g.cc()[0], oracle.cc(ptr, to)[0])
    ms_src = np.array(_sources(ptr, len(to), 3, 17), np.uint32)
    if graph == "chain":  # thousands of levels: past the u8 levels of the batched search
        with pytest.raises(og.OgError) as e:
            g.msbfs(ms_src)
        assert e.value.name == "OG_EUNSUPPORTED" and "254 levels" in str(e.value)
    else:
        lvm, _, _, _ = g.msbfs(ms_src)
        assert np.array_equal(lvm[0], oracle.bfs(ptr, to, int(ms_src[0]))[0])
    g.close()


def test_direction_optimising_needs_hbm_resident_graph(gpu, rmat_w):
    with pytest.raises(og.OgError) as e:
        _prepare(rmat_w, "ascetic+diropt", 0.25, 0.8, 8)
    assert e.value.name == "OG_EUNSUPPORTED"


def test_probe_host_link(gpu):
    """The link probe runs its three measurements and reports rates."""
    r = og.probe_host_link(1 << 22)
    assert set(r) == {"contiguous_gbs", "scattered_rows_gbs", "memcpy_gbs"}
    assert all(v > 0 for v in r.values())
    with pytest.raises(og.OgError):
        og.probe_host_link(1 << 10)


@pytest.mark.parametrize("approach,frac,k", [("ascetic", 1.0, 1.0), ("ascetic", 0.25, 0.8), ("ascetic", 1 / 64, 0.5),
                                              ("ascetic+zerocopy", 0.25, 0.8), ("ascetic+zerocopy", 1 / 64, 0.5),
                                              ("subway+zerocopy", 0.25, 0.0), ("subway", 1 / 64, 0.0),
                                              ("ascetic+zerocopy+nograph", 0.25, 0.8),
                                              ("ascetic+zerocopy+nofuse", 0.25, 0.8)])
@pytest.mark.parametrize("split,delta", [(32, 32), (8, 32), (40, 16), (16, 0)])
def test_light_heavy_split(gpu, rmat_w, approach, frac, k, split, delta):
    """Rows stored light edges first: near-far SSSP relaxes light edges and
    each vertex's heavy edges once its bucket settled; any split weight /
    bucket width pair (and delta 0, which ignores the split) gives the
    oracle's distances, and BFS / CC / multi-source BFS on the same
    preparation use whole rows."""
    ptr, to, w, _ = rmat_w.export()
    E = len(to)
    _say("light/heavy split rmat13", approach, frac, k, split, delta)
    name, *opts = approach.split("+")
    flags = 0
    for o in opts:
        flags |= {"zerocopy": og.PLAN_ZERO_COPY, "nograph": og.PLAN_NO_GRAPH, "nofuse": og.PLAN_NO_FUSE}[o]
    rmat_w.prepare(name, edge_budget_bytes=int(max(1, E * 4 * frac)), advisek=k, flags=flags, split_weight=split)
    rmat_w.set_tuning(delta=delta)
    for s in _sources(ptr, E, 3, 17):
        d, st = rmat_w.sssp(s)
        assert np.array_equal(d, oracle.sssp(ptr, to, w, s)), s
        assert st["checksum"] == og.levels_checksum(d)
    assert np.array_equal(rmat_w.bfs(5)[0], oracle.bfs(ptr, to, 5)[0])
    assert np.array_equal(rmat_w.cc()[0], oracle.cc(ptr, to)[0])
    lv, _, _, _ = rmat_w.msbfs([5, 9])
    assert np.array_equal(lv[1], oracle.bfs(ptr, to, 9)[0])
    rmat_w.set_tuning(delta=0)


def test_light_heavy_split_refused(gpu):
    """The split needs the packed weighted format."""
    g = og.Graph.rmat(10, 8, seed=3, weight_seed=4)
    with pytest.raises(og.OgError) as e:
        g.prepare("ascetic", edge_budget_bytes=g.num_edges * 8, advisek=1.0, flags=og.PLAN_NO_PACK, split_weight=16)
    assert e.value.name == "OG_EUNSUPPORTED"
    h = og.Graph.rmat(10, 8, seed=3)
    with pytest.raises(og.OgError):
        h.prepare("ascetic", edge_budget_bytes=h.num_edges * 4, advisek=1.0, split_weight=16)
    g.close()
    h.close()
