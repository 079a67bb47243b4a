"""Continuous-batching server: each query's result is the same alone and in
the mix (BFS slots batched 64 to an engine; SSSP queries on single-source
lanes, or on request batched 16 to an engine with their own near-far
thresholds), equals the single-query path and the oracle; admission,
reservation and failure rules hold."""
import sys
import time

import numpy as np
import pytest

import onegraph_amd as og
import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=["topdown", "diropt", "diropt+split"])
def graphs(request):
    """Three graphs; "diropt" prepares them with their in-edges, so BFS slots
    with dense frontiers run bottom-up rounds (graphs.in_edges tells which);
    "+split" also stores the weighted graphs' rows light edges first, so the
    SSSP lanes run delta-stepping with wide buckets while the BFS slots read
    whole rows."""
    gs = [og.Graph.rmat(12, 16, seed=51, weight_seed=52), og.Graph.er(3000, 8, seed=53).add_weights(54),
          og.Graph.rmat(11, 8, seed=55)]
    flags = og.PLAN_DIRECTION_OPT if request.param.startswith("diropt") else 0
    split = request.param.endswith("+split")
    for g in gs:
        weighted = g.info()["kind"] == og.KIND_WEIGHTED
        g.prepare("ascetic", edge_budget_bytes=g.num_edges * (8 if weighted else 4), advisek=1.0, flags=flags,
                  split_weight=8 if split and weighted else 0)
        if split and weighted:
            g.set_tuning(delta=256)
    yield _Graphs(gs, request.param.startswith("diropt"))
    for g in gs:
        g.close()


class _Graphs(list):
    def __init__(self, gs, in_edges):
        super().__init__(gs)
        self.in_edges = in_edges


def _state_bytes(graphs):
    """Admission rule of og_server_create: per graph the BFS engine's masks
    (24 V bytes) and 4 kept-levels buffers (16 V), per weighted graph the
    SSSP engine's 16 distance columns (64 V), light and heavy pending masks
    (8 V), active list and masks (8 V), work prefix (8 V + 8), extraction
    buffer (4 V) and
    the relax's chunk table (4 (16 E / 1024 + 2))."""
    total = 0
    for g in graphs:
        V = g.num_vertices
        total += 40 * V
        if g.info()["kind"] == og.KIND_WEIGHTED:
            total += 92 * V + 8 + 4 * (16 * g.num_edges // 1024 + 2)
    return total


def _queries(graphs, n, seed):
    rng = np.random.default_rng(seed)
    qs = []
    for _ in range(n):
        gi = int(rng.integers(len(graphs)))
        g = graphs[gi]
        weighted = g.info()["kind"] == og.KIND_WEIGHTED
        kind = og.QUERY_SSSP if (weighted and rng.random() < 0.3) else og.QUERY_BFS
        ptr, to, _, _ = g.export()
        deg = np.diff(np.append(ptr.astype(np.int64), len(to)))
        qs.append((gi, kind, int(rng.choice(np.flatnonzero(deg > 0)))))
    return qs


def _wait_all(srv, ids, timeout=60):
    got = {}
    t0 = time.time()
    while len(got) < len(ids):
        for r in srv.poll(256, timeout_ms=200):
            got[r["id"]] = r
        assert time.time() - t0 < timeout, "server did not finish"
    return got


def _oracle(graphs, gi, kind, src):
    ptr, to, w, _ = graphs[gi].export()
    return oracle.bfs(ptr, to, src)[0] if kind == og.QUERY_BFS else oracle.sssp(ptr, to, w, src)


def test_alone_and_in_the_mix(gpu, graphs):
    qs = _queries(graphs, 60, 7)
    _say("server mix of", len(qs))
    # the exact budget: queries join the BFS and SSSP batches mid-flight;
    # kept BFS levels use 4 buffers per graph, so those queries queue
    # the batched engine alone (no lanes), at the exact budget
    with og.Server(graphs, state_budget_bytes=_state_bytes(graphs), flags=og.SERVER_NO_SSSP_LANES) as srv:
        info = srv.info()
        assert info["bfs_slots"] == [64, 64, 64] and info["sssp_slots"] == [16, 16, 0]
        ids = [srv.submit(gi, k, s, keep=True) for gi, k, s in qs]
        res = _wait_all(srv, ids)
        mix = {i: srv.take_values(i, qs[n][0]) for n, i in enumerate(ids)}
        alone = {}
        for n, (gi, k, s) in enumerate(qs[:20]):
            i = srv.submit(gi, k, s, keep=True)
            r = _wait_all(srv, [i])[i]
            assert r["status"] == 0
            alone[n] = (r, srv.take_values(i, gi))
        info = srv.info()
        assert info["completed"] == len(qs) + 20 and info["in_flight"] == 0
    for n, (gi, k, s) in enumerate(qs):
        r = res[ids[n]]
        assert r["status"] == 0
        ref = _oracle(graphs, gi, k, s)
        assert np.array_equal(mix[ids[n]], ref), (n, gi, k, s)
        assert r["checksum"] == og.levels_checksum(ref)
        assert r["reached"] == int((ref != og.UNREACHED).sum())
        if n in alone:
            assert np.array_equal(alone[n][1], mix[ids[n]])
            assert alone[n][0]["checksum"] == r["checksum"]


def test_matches_single_query_path(gpu, graphs):
    qs = _queries(graphs, 30, 9)
    with og.Server(graphs) as srv:
        ids = [srv.submit(gi, k, s) for gi, k, s in qs]
        res = _wait_all(srv, ids)
    for n, (gi, k, s) in enumerate(qs):  # graphs are free again after destroy
        g = graphs[gi]
        _, st = (g.bfs(s, want_values=False) if k == og.QUERY_BFS else g.sssp(s, want_values=False))
        assert st["checksum"] == res[ids[n]]["checksum"]


def test_admission_and_reservation(gpu, graphs):
    g = graphs[0]
    with pytest.raises(og.OgError) as e:
        og.Server([g], state_budget_bytes=16)
    assert e.value.name == "OG_EBUDGET"
    with pytest.raises(og.OgError) as e:  # one byte short
        og.Server([g], state_budget_bytes=_state_bytes([g]) - 1, flags=og.SERVER_NO_SSSP_LANES)
    assert e.value.name == "OG_EBUDGET"
    with pytest.raises(og.OgError) as e:  # BFS state only: no SSSP lane fits
        og.Server([g], state_budget_bytes=40 * g.num_vertices)
    assert e.value.name == "OG_EBUDGET"
    srv = og.Server([g], state_budget_bytes=_state_bytes([g]), max_pending=4, flags=og.SERVER_NO_SSSP_LANES)
    assert srv.info()["bfs_slots"] == [64] and srv.info()["sssp_slots"] == [16]
    assert srv.info()["state_bytes"] == _state_bytes([g])
    with pytest.raises(og.OgError) as e:  # reserved while served
        g.bfs(0)
    assert e.value.name == "OG_EBUSY"
    with pytest.raises(og.OgError) as e:
        srv.submit(0, og.QUERY_BFS, g.num_vertices)
    assert e.value.name == "OG_EINVAL"
    rejected = 0
    for _ in range(200):
        try:
            srv.submit(0, og.QUERY_SSSP, 0)
        except og.OgError as ex:
            assert ex.name == "OG_EBUSY"
            rejected += 1
    assert rejected > 0 and srv.info()["rejected"] == rejected
    srv.close()
    g.bfs(0)  # released
    # a graph whose edges are not all on the device is refused
    h = og.Graph.rmat(10, 8, seed=3)
    h.prepare("ascetic", edge_budget_bytes=h.num_edges, advisek=0.5)
    with pytest.raises(og.OgError) as e:
        og.Server([h])
    assert e.value.name == "OG_EUNSUPPORTED"
    h.close()


def test_burst_of_full_batches(gpu, graphs):
    """64 BFS queries arriving at once fill every slot of an engine (their
    dense rounds pull or push all together); queries arriving one by one
    afterwards run alone.  Every checksum equals the plain path's."""
    g = graphs[0]
    ptr, to, _, _ = g.export()
    deg = np.diff(np.append(ptr.astype(np.int64), len(to)))
    srcs = np.random.default_rng(17).choice(np.flatnonzero(deg > 0), 96, replace=False)
    _say("burst of 64 BFS, then 32 one by one, in_edges", graphs.in_edges)
    with og.Server([g]) as srv:
        ids = [srv.submit(0, og.QUERY_BFS, int(s)) for s in srcs[:64]]
        res = _wait_all(srv, ids)
        for s in srcs[64:]:
            i = srv.submit(0, og.QUERY_BFS, int(s))
            ids.append(i)
            res.update(_wait_all(srv, [i]))
            time.sleep(0.001)
    for n, s in enumerate(srcs):
        _, st = g.bfs(int(s), want_values=False)
        assert st["checksum"] == res[ids[n]]["checksum"], (n, int(s))


def test_sssp_batches_alone_and_in_the_mix(gpu, graphs):
    """More SSSP queries than slots at once (repeated sources among them),
    then one at a time: each query's distances equal the oracle's and its
    result alone, whatever else shares its rounds."""
    gi = 0
    g = graphs[gi]
    ptr, to, w, _ = g.export()
    deg = np.diff(np.append(ptr.astype(np.int64), len(to)))
    cand = np.flatnonzero(deg > 0)
    srcs = [int(x) for x in np.random.default_rng(19).choice(cand, 36, replace=False)] + [int(cand[0])] * 3
    _say("sssp burst of", len(srcs), "on graph 0")
    with og.Server([g, graphs[1]], flags=og.SERVER_SSSP_BATCH) as srv:  # lanes, then batches
        ids = [srv.submit(0, og.QUERY_SSSP, s, keep=True) for s in srcs]
        ids += [srv.submit(1, og.QUERY_SSSP, 5, keep=True)]  # the other engine runs beside it
        res = _wait_all(srv, ids)
        vals = {i: srv.take_values(i, 0 if n < len(srcs) else 1) for n, i in enumerate(ids)}
        alone = {}
        for s in srcs[:4]:
            i = srv.submit(0, og.QUERY_SSSP, s, keep=True)
            alone[s] = (_wait_all(srv, [i])[i], srv.take_values(i, 0))
    for n, s in enumerate(srcs):
        r = res[ids[n]]
        assert r["status"] == 0
        ref = oracle.sssp(ptr, to, w, s)
        assert np.array_equal(vals[ids[n]], ref), (n, s)
        assert r["checksum"] == og.levels_checksum(ref)
        assert r["reached"] == int((ref != og.UNREACHED).sum())
        assert r["traversed_edges"] == int(deg[ref != og.UNREACHED].sum())
        if s in alone:
            assert alone[s][0]["checksum"] == r["checksum"]
            assert np.array_equal(alone[s][1], vals[ids[n]])
    p1, t1, w1, _ = graphs[1].export()
    assert np.array_equal(vals[ids[-1]], oracle.sssp(p1, t1, w1, 5))


@pytest.mark.parametrize("delta", [0, 1, 32])
def test_sssp_batch_bucket_widths(gpu, delta):
    """Per-slot thresholds at every bucket width, including none (delta 0:
    every pending entry relaxes each round) and width 1 (a threshold move
    for nearly every distance)."""
    g = og.Graph.rmat(11, 8, seed=61, weight_seed=62)
    g.prepare("ascetic", edge_budget_bytes=g.num_edges * 8, advisek=1.0)
    g.set_tuning(delta=delta)
    ptr, to, w, _ = g.export()
    cand = np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), len(to))) > 0)
    srcs = [int(x) for x in np.random.default_rng(63).choice(cand, 20, replace=False)]
    lanes_off = og.SERVER_NO_SSSP_LANES  # the batched engine takes every query
    with og.Server([g], flags=lanes_off) as srv:
        ids = [srv.submit(0, og.QUERY_SSSP, s) for s in srcs]
        res = _wait_all(srv, ids)
    with og.Server([g], sssp_delta=7, flags=lanes_off) as srv:  # the server's own width over the graph's
        ids2 = [srv.submit(0, og.QUERY_SSSP, s) for s in srcs]
        res2 = _wait_all(srv, ids2)
    g.prepare("ascetic", edge_budget_bytes=g.num_edges * 8, advisek=1.0, split_weight=8)
    g.set_tuning(delta=delta)
    with og.Server([g], sssp_delta=256 if delta == 32 else 0, flags=lanes_off) as srv:  # light / heavy rounds
        ids3 = [srv.submit(0, og.QUERY_SSSP, s) for s in srcs[:10]]
        time.sleep(0.002)
        ids3 += [srv.submit(0, og.QUERY_SSSP, s) for s in srcs[10:]]  # joining running slots
        res3 = _wait_all(srv, ids3)
    for n, s in enumerate(srcs):
        want = og.levels_checksum(oracle.sssp(ptr, to, w, s))
        assert res[ids[n]]["checksum"] == want, (delta, s)
        assert res2[ids2[n]]["checksum"] == want, (7, s)
        assert res3[ids3[n]]["checksum"] == want, ("split", s)
    g.close()


def test_failed_engine_refuses_its_queue(gpu, graphs):
    """An engine that stops (forced here from the host: the SSSP engines of a
    server created with og.SERVER_TEST_FAIL_SSSP raise at their first round,
    as after a device error, without touching the GPU) fails the queries it
    held and every later submit to it; the BFS engines keep serving."""
    g = graphs[0]
    for extra in (0, og.SERVER_NO_SSSP_LANES):  # a lane fails; the batched engine fails
        with og.Server([g], flags=og.SERVER_TEST_FAIL_SSSP | extra) as srv:
            i = srv.submit(0, og.QUERY_SSSP, 0)
            r = _wait_all(srv, [i])[i]
            assert og.STATUS[r["status"]] == "OG_EHIP"
            with pytest.raises(og.OgError) as e:
                srv.submit(0, og.QUERY_SSSP, 1)
            assert e.value.name == "OG_EHIP" and "stopped" in str(e.value)
            j = srv.submit(0, og.QUERY_BFS, 0)
            rb = _wait_all(srv, [j])[j]
            assert rb["status"] == 0
            assert srv.info()["in_flight"] == 0 and srv.info()["pending"] == 0
    _, st = g.bfs(0, want_values=False)
    assert st["checksum"] == rb["checksum"]


@pytest.mark.parametrize("split", [0, 8])
def test_sssp_batch_many_sweep_blocks(gpu, split):
    """A graph spanning several select / write workgroups (V = 2^16: 8 blocks
    of 8192 vertices), with and without light/heavy rows, 24 queries."""
    g = og.Graph.rmat(16, 8, seed=71, weight_seed=72)
    g.prepare("ascetic", edge_budget_bytes=g.num_edges * 8, advisek=1.0, split_weight=split)
    ptr, to, w, _ = g.export()
    cand = np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), len(to))) > 0)
    srcs = [int(x) for x in np.random.default_rng(73).choice(cand, 24, replace=False)]
    _say("sssp batch rmat16 split", split)
    with og.Server([g], sssp_delta=256 if split else 64, flags=og.SERVER_NO_SSSP_LANES) as srv:
        ids = [srv.submit(0, og.QUERY_SSSP, s) for s in srcs]
        res = _wait_all(srv, ids)
    for n, s in enumerate(srcs):
        assert res[ids[n]]["checksum"] == og.levels_checksum(oracle.sssp(ptr, to, w, s)), (split, s)
    g.close()


def test_sssp_lanes_first(gpu, graphs):
    """SSSP queries run on single-source lanes by default (a lone query and a
    burst alike); with og.SERVER_SSSP_BATCH a burst fills both lanes and then
    the batched engine (8 or more waiting), while fewer waiting queries wait
    for a lane.  The budget decides how many lanes a weighted graph gets."""
    g = graphs[0]
    ptr, to, w, _ = g.export()
    deg = np.diff(np.append(ptr.astype(np.int64), len(to)))
    cand = np.flatnonzero(deg > 0)
    srcs = [int(x) for x in np.random.default_rng(29).choice(cand, 40, replace=False)]
    _, st0 = g.sssp(srcs[0], want_values=False)
    bfs_bytes = 40 * g.num_vertices

    def check(srv, burst, paths_want):
        ids = [srv.submit(0, og.QUERY_SSSP, s, keep=True) for s in burst]
        res = _wait_all(srv, ids)
        paths = [res[i]["path"] for i in ids]
        assert paths_want(paths), paths
        for n, s in enumerate(burst):
            ref = oracle.sssp(ptr, to, w, s)
            assert res[ids[n]]["status"] == 0
            assert np.array_equal(srv.take_values(ids[n], 0), ref), s
            assert res[ids[n]]["checksum"] == og.levels_checksum(ref)
            assert res[ids[n]]["traversed_edges"] == int(deg[ref != og.UNREACHED].sum())

    with og.Server([g]) as srv:  # default: two lanes, no batched engine
        assert srv.info()["sssp_slots"] == [2]
        lanes2 = srv.info()["state_bytes"] - bfs_bytes
        i = srv.submit(0, og.QUERY_SSSP, srcs[0], keep=True)
        r = _wait_all(srv, [i])[i]
        assert r["status"] == 0 and r["path"] == og.PATH_LANE
        assert np.array_equal(srv.take_values(i, 0), oracle.sssp(ptr, to, w, srcs[0]))
        assert r["checksum"] == st0["checksum"]
        _say("lanes only: burst of 35")
        check(srv, srcs[5:40], lambda p: set(p) == {og.PATH_LANE})
        info = srv.info()
        assert info["in_flight"] == 0 and info["pending"] == 0 and info["completed"] == 36
    with og.Server([g], flags=og.SERVER_SSSP_BATCH) as srv:  # lanes, then the batched engine
        assert srv.info()["sssp_slots"] == [16 + 2]
        assert srv.info()["state_bytes"] == _state_bytes([g]) + lanes2
        _say("lanes + batch: bursts of 4 and 35")
        check(srv, srcs[1:5], lambda p: set(p) == {og.PATH_LANE})  # never 8 waiting
        check(srv, srcs[5:40], lambda p: og.PATH_BATCH in p and p.count(og.PATH_LANE) >= 2)
    # the BFS state plus one lane's bytes: one lane
    with og.Server([g], state_budget_bytes=bfs_bytes + lanes2 // 2) as srv:
        assert srv.info()["sssp_slots"] == [1]
        assert srv.info()["state_bytes"] == bfs_bytes + lanes2 // 2
        check(srv, srcs[:12], lambda p: set(p) == {og.PATH_LANE})
    with og.Server([g], flags=og.SERVER_NO_SSSP_LANES) as srv:
        assert srv.info()["sssp_slots"] == [16]
        assert srv.info()["state_bytes"] == _state_bytes([g])
    with og.Server([g], sssp_lanes=4) as srv:
        assert srv.info()["sssp_slots"] == [4]
        assert srv.info()["state_bytes"] == bfs_bytes + 2 * lanes2
        check(srv, srcs[:11], lambda p: set(p) == {og.PATH_LANE})
    with pytest.raises(og.OgError) as e:
        og.Server([g], sssp_lanes=9)
    assert e.value.name == "OG_EINVAL"


def test_sssp_running_cap(gpu, graphs):
    """With at most one SSSP lane running at once over both weighted graphs,
    bursts on both still all complete with the oracle's distances, and BFS
    queries beside them are served."""
    rng = np.random.default_rng(31)
    qs = []
    for gi in (0, 1):
        ptr, to, w, _ = graphs[gi].export()
        cand = np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), len(to))) > 0)
        qs += [(gi, int(s)) for s in rng.choice(cand, 10, replace=False)]
    with og.Server([graphs[0], graphs[1]], sssp_running_max=1) as srv:
        assert srv.info()["sssp_slots"] == [2, 2]
        ids = [srv.submit(gi, og.QUERY_SSSP, s) for gi, s in qs]
        bfs = [srv.submit(gi, og.QUERY_BFS, s) for gi, s in qs[::4]]
        res = _wait_all(srv, ids + bfs)
        assert srv.info()["in_flight"] == 0 and srv.info()["pending"] == 0
    for n, (gi, s) in enumerate(qs):
        ptr, to, w, _ = graphs[gi].export()
        assert res[ids[n]]["checksum"] == og.levels_checksum(oracle.sssp(ptr, to, w, s)), (gi, s)
        assert res[ids[n]]["path"] == og.PATH_LANE
    for n, (gi, s) in enumerate(qs[::4]):
        _, st = graphs[gi].bfs(s, want_values=False)
        assert res[bfs[n]]["checksum"] == st["checksum"]


def test_deadline_expires_queued_queries(gpu):
    """With a 1 ms admission deadline, a burst of 3000 BFS queries (64 slots)
    leaves most of them queued past it: they complete with OG_ETIMEDOUT, the
    admitted ones with their usual results, and nothing is left pending.
    Without a deadline none expire.  The graph (R-MAT 18, 4 M edges) makes a
    batch of 64 traversals take longer than submitting 64 queries, so the
    queue grows whatever the speed of the GPU (on a graph of a few thousand
    vertices the engine kept pace with the submits and nothing expired)."""
    g = og.Graph.rmat(18, 16, seed=57)
    g.prepare("ascetic", edge_budget_bytes=g.num_edges * 4, advisek=1.0)
    ptr, to, _, _ = g.export()
    deg = np.diff(np.append(ptr.astype(np.int64), len(to)))
    srcs = np.random.default_rng(23).choice(np.flatnonzero(deg > 0), 3000)
    with og.Server([g], deadline_ms=1) as srv:
        ids = [srv.submit(0, og.QUERY_BFS, int(s)) for s in srcs]
        res = _wait_all(srv, ids, timeout=120)
        info = srv.info()
        assert info["pending"] == 0 and info["in_flight"] == 0 and info["completed"] == len(ids)
    names = [og.STATUS[res[i]["status"]] for i in ids]
    assert set(names) <= {"OG_OK", "OG_ETIMEDOUT"}
    assert names.count("OG_ETIMEDOUT") > 0 and names.count("OG_OK") >= 64
    ok = [n for n, i in enumerate(ids) if res[i]["status"] == 0][:20]
    for n in ok:
        _, st = g.bfs(int(srcs[n]), want_values=False)
        assert st["checksum"] == res[ids[n]]["checksum"]
    with og.Server([g]) as srv:
        ids = [srv.submit(0, og.QUERY_BFS, int(s)) for s in srcs[:500]]
        res = _wait_all(srv, ids, timeout=120)
    assert all(res[i]["status"] == 0 for i in ids)
    g.close()
