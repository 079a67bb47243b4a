"""Pins the CPU oracle: golden fixtures from scipy/networkx, closed forms,
invariants, and the reference's own serial CC program (oracle/_ref)."""
import glob
import os
import subprocess

import numpy as np
import pytest

import oracle

U = 0xFFFFFFFF
GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "graph_*.npz")))
REF_CC = os.path.join(os.path.dirname(oracle.__file__), "_ref", "cc_serial")


def csr_from_edges(V, edges, weights=None):
    edges = sorted(edges) if weights is None else edges
    src = np.array([e[0] for e in edges], np.int64)
    dst = np.array([e[1] for e in edges], np.uint32)
    counts = np.bincount(src, minlength=V) if len(src) else np.zeros(V, np.int64)
    ptr = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint64)
    return ptr, dst, (np.array(weights, np.uint32) if weights is not None else None)


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p))
def test_golden(path):
    z = np.load(path)
    ptr, cols, w = z["ptr"], z["cols"], z["weights"]
    for i, s in enumerate(z["sources"]):
        lv, _, _ = oracle.bfs(ptr, cols, int(s))
        assert np.array_equal(lv, z["bfs"][i])
        assert np.array_equal(oracle.sssp(ptr, cols, w, int(s)), z["sssp"][i])
    lab, _ = oracle.cc(ptr, cols)
    assert np.array_equal(lab, z["cc"])
    # PageRank over the pull layout of the same graph
    V, E = len(ptr), len(cols)
    src = np.repeat(np.arange(V), np.diff(np.append(ptr.astype(np.int64), E)))
    order = np.lexsort((src, cols))
    in_src = src[order].astype(np.uint32)
    in_ptr = np.concatenate([[0], np.cumsum(np.bincount(cols, minlength=V))[:-1]]).astype(np.uint64)
    outdeg = np.bincount(src, minlength=V).astype(np.uint32)
    r, it, l1 = oracle.pagerank(outdeg, in_ptr, in_src, 0.85, 1e-15, 10000)
    assert np.abs(r - z["pagerank"]).max() < 1e-12


def test_bfs_closed_forms():
    n = 50
    ptr, cols, _ = csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])
    lv, ns, it = oracle.bfs(ptr, cols, 0)
    assert lv.tolist() == list(range(1, n + 1)) and ns == n and it == n
    lv, _, _ = oracle.bfs(ptr, cols, 10)
    assert lv.tolist() == [U] * 10 + list(range(1, n - 10 + 1))
    # star: center 0 -> leaves
    ptr, cols, _ = csr_from_edges(9, [(0, i) for i in range(1, 9)])
    assert oracle.bfs(ptr, cols, 0)[0].tolist() == [1] + [2] * 8
    assert oracle.bfs(ptr, cols, 3)[0].tolist() == [U, U, U, 1, U, U, U, U, U]
    # complete digraph
    k = 7
    ptr, cols, _ = csr_from_edges(k, [(a, b) for a in range(k) for b in range(k) if a != b])
    assert oracle.bfs(ptr, cols, 4)[0].tolist() == [2, 2, 2, 2, 1, 2, 2]
    # complete binary tree: level = floor(log2(i+1)) + 1
    m = 127
    ptr, cols, _ = csr_from_edges(m, [(i, c) for i in range(m) for c in (2 * i + 1, 2 * i + 2) if c < m])
    assert oracle.bfs(ptr, cols, 0)[0].tolist() == [int(np.floor(np.log2(i + 1))) + 1 for i in range(m)]
    # 2-D grid with edges both ways: Manhattan distance + 1
    R, Cc = 6, 9
    vid = lambda r, c: r * Cc + c
    ed = []
    for r in range(R):
        for c in range(Cc):
            for dr, dc in ((0, 1), (1, 0), (0, -1), (-1, 0)):
                if 0 <= r + dr < R and 0 <= c + dc < Cc:
                    ed.append((vid(r, c), vid(r + dr, c + dc)))
    ptr, cols, _ = csr_from_edges(R * Cc, ed)
    lv = oracle.bfs(ptr, cols, vid(2, 3))[0]
    assert lv.tolist() == [abs(r - 2) + abs(c - 3) + 1 for r in range(R) for c in range(Cc)]


def test_bfs_self_loops_and_empty():
    ptr, cols, _ = csr_from_edges(3, [(0, 0), (0, 1), (1, 1)])
    assert oracle.bfs(ptr, cols, 0)[0].tolist() == [1, 2, U]
    ptr = np.zeros(4, np.uint64)
    lv, ns, it = oracle.bfs(ptr, np.zeros(0, np.uint32), 2)
    assert lv.tolist() == [U, U, 1, U] and ns == 1 and it == 1


def test_sssp_closed_forms_and_invariants():
    # path with weights: prefix sums + 1
    w = [5, 1, 63, 2, 7]
    ptr, cols, ww = csr_from_edges(6, [(i, i + 1) for i in range(5)], w)
    assert oracle.sssp(ptr, cols, ww, 0).tolist() == [1, 6, 7, 70, 72, 79]
    # unit weights reduce to BFS
    rng = np.random.default_rng(3)
    V = 400
    ed = sorted(set((int(a), int(b)) for a, b in rng.integers(0, V, (2000, 2)) if a != b))
    ptr, cols, _ = csr_from_edges(V, ed)
    assert np.array_equal(oracle.sssp(ptr, cols, np.ones(len(cols), np.uint32), 7), oracle.bfs(ptr, cols, 7)[0])
    # triangle inequality on every edge, tightness through a predecessor
    wt = rng.integers(1, 64, len(cols)).astype(np.uint32)
    d = oracle.sssp(ptr, cols, wt, 7).astype(np.int64)
    src = np.repeat(np.arange(V), np.diff(np.append(ptr.astype(np.int64), len(cols))))
    reached = d[src] != U
    assert np.all(d[cols[reached]] <= d[src[reached]] + wt[reached])
    tight = np.zeros(V, bool)
    tight[cols[reached][d[cols[reached]] == d[src[reached]] + wt[reached]]] = True
    assert np.all(tight[(d != U) & (np.arange(V) != 7)])


def test_cc_closed_forms():
    # directed path: every vertex is reached from 0
    ptr, cols, _ = csr_from_edges(5, [(i, i + 1) for i in range(4)])
    assert oracle.cc(ptr, cols)[0].tolist() == [0, 0, 0, 0, 0]
    # reversed path: nothing flows to higher labels... 4->3->2->1->0
    ptr, cols, _ = csr_from_edges(5, [(i + 1, i) for i in range(4)])
    assert oracle.cc(ptr, cols)[0].tolist() == [0, 1, 2, 3, 4]
    # two undirected components {0,3,5} {1,2,4}
    und = [(0, 3), (3, 0), (3, 5), (5, 3), (1, 4), (4, 1), (4, 2), (2, 4)]
    ptr, cols, _ = csr_from_edges(6, und)
    assert oracle.cc(ptr, cols)[0].tolist() == [0, 1, 1, 0, 1, 0]


def test_cc_union_find_random():
    rng = np.random.default_rng(5)
    V = 3000
    pairs = rng.integers(0, V, (2500, 2))
    parent = list(range(V))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    ed = set()
    for a, b in pairs:
        a, b = int(a), int(b)
        if a != b:
            ed.add((a, b))
            ed.add((b, a))
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
    mins = {}
    for v in range(V):
        r = find(v)
        mins[r] = min(mins.get(r, v), v)
    ptr, cols, _ = csr_from_edges(V, sorted(ed))
    assert oracle.cc(ptr, cols)[0].tolist() == [mins[find(v)] for v in range(V)]


@pytest.mark.skipif(not os.path.exists(REF_CC), reason="reference serial CC not built (needs /root/reference)")
def test_cc_update_count_matches_reference_program(tmp_path):
    """The reference CC program prints the number of successful lowerings of
    its Gauss-Seidel sweep (CC/serial/serial.cpp:195,225); the oracle must
    count the same on the same file."""
    import onegraph_amd as og

    for scale, ef, seed in [(8, 8, 1), (11, 8, 2), (12, 4, 3)]:
        g = og.Graph.rmat(scale, ef, seed=seed)
        f = str(tmp_path / "g.bcsr")
        g.save(f)
        ptr, cols, _, _ = g.export()
        _, n = oracle.cc(ptr, cols)
        out = subprocess.run([REF_CC, "-f", f], capture_output=True, text=True, timeout=60).stdout
        ref_n = int([l for l in out.splitlines() if l.startswith("nodeSum")][0].split(":")[1])
        assert n == ref_n
        g.close()


def test_pagerank_closed_forms():
    d = 0.85
    # directed cycle: uniform
    n = 10
    ptr = np.arange(n, dtype=np.uint64)
    in_cols = np.array([(i - 1) % n for i in range(n)], np.uint32)
    r, it, l1 = oracle.pagerank(np.ones(n, np.uint32), ptr, in_cols, d, 1e-14, 1000)
    assert np.abs(r - 1 / n).max() < 1e-15
    # complete digraph: uniform
    k = 6
    od = np.full(k, k - 1, np.uint32)
    ptr = np.arange(0, k * (k - 1), k - 1, dtype=np.uint64)
    in_cols = np.array([u for v in range(k) for u in range(k) if u != v], np.uint32)
    r, _, _ = oracle.pagerank(od, ptr, in_cols, d, 1e-14, 1000)
    assert np.abs(r - 1 / k).max() < 1e-15
    # undirected star, center 0 with m leaves
    m = 9
    nn = m + 1
    od = np.array([m] + [1] * m, np.uint32)
    ptr = np.array([0] + [m + i for i in range(m)], np.uint64)
    in_cols = np.array(list(range(1, m + 1)) + [0] * m, np.uint32)
    r, _, _ = oracle.pagerank(od, ptr, in_cols, d, 1e-15, 10000)
    leaf = (1 + d / m) / (nn * (1 + d))
    center = (1 - d) / nn + d * m * leaf
    assert abs(r[0] - center) < 1e-13 and np.abs(r[1:] - leaf).max() < 1e-13
    # all dangling: stays uniform; mass conserved
    r, it, _ = oracle.pagerank(np.zeros(4, np.uint32), np.zeros(4, np.uint64), np.zeros(0, np.uint32), d, 1e-9, 50)
    assert np.abs(r - 0.25).max() < 1e-15 and it == 1


def _pull_of(ptr, cols):
    """Pull layout (out-degree, in-edge CSR sorted by source) of a CSR."""
    V, E = len(ptr), len(cols)
    src = np.repeat(np.arange(V), np.diff(np.append(ptr.astype(np.int64), E)))
    order = np.lexsort((src, cols))
    in_src = src[order].astype(np.uint32)
    in_ptr = np.concatenate([[0], np.cumsum(np.bincount(cols, minlength=V))[:-1]]).astype(np.uint64)
    outdeg = np.bincount(src, minlength=V).astype(np.uint32)
    return outdeg, in_ptr, in_src


def _er_pull(n, avg, seed):
    """Symmetrised G(n, m = n * avg / 2) in numpy, as a pull layout."""
    rng = np.random.default_rng(seed)
    m = n * avg // 2
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    keep = a != b
    u = np.concatenate([a[keep], b[keep]])
    v = np.concatenate([b[keep], a[keep]])
    key = np.unique(u.astype(np.int64) * n + v)
    u, v = key // n, (key % n).astype(np.uint32)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(u, minlength=n))[:-1]]).astype(np.uint64)
    return _pull_of(ptr, v)


def _rmat_like(n, m, seed):
    """A skewed directed graph with many dangling vertices (no out-edges)."""
    rng = np.random.default_rng(seed)
    a = (rng.pareto(1.2, m) * 3).astype(np.int64) % n
    b = rng.integers(0, n, m)
    keep = a != b
    key = np.unique(a[keep] * n + b[keep])
    u, v = key // n, (key % n).astype(np.uint32)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(u, minlength=n))[:-1]]).astype(np.uint64)
    return _pull_of(ptr, v)


def _pull_cases():
    cases = []
    for path in GOLDEN:
        z = np.load(path)
        cases.append((os.path.basename(path), _pull_of(z["ptr"], z["cols"])))
    cases.append(("er_2^16_avg32", _er_pull(1 << 16, 32, 11)))
    cases.append(("skewed_dangling", _rmat_like(20000, 120000, 12)))
    return cases


@pytest.mark.parametrize("threads", [1, 2, 16])
def test_pagerank_par_equals_serial(threads):
    """og_oracle_pagerank_par (the CPU baseline's all-core variant) against
    the pinned serial function: ranks bit for bit, same iteration count and
    residual, including graphs with dangling mass and ones that hit the
    iteration cap (tolerance 0)."""
    for name, (od, ip, ic) in _pull_cases():
        for tol, cap in ((1e-7, 100), (0.0, 7)):
            r0, it0, l0 = oracle.pagerank(od, ip, ic, 0.85, tol, cap, parallel=False)
            r1, it1, l1 = oracle.pagerank(od, ip, ic, 0.85, tol, cap, threads=threads, parallel=True)
            assert it0 == it1, name
            assert l0 == l1, name
            assert np.array_equal(r0.view(np.uint64), r1.view(np.uint64)), name
        assert abs(r0.sum() - 1.0) < 1e-9, name


def test_pagerank_par_on_closed_form():
    """The threaded function on its own: a directed cycle stays uniform and
    the star's closed form holds at 16 threads."""
    n = 10000
    ptr = np.arange(n, dtype=np.uint64)
    in_cols = np.array([(i - 1) % n for i in range(n)], np.uint32)
    r, _, _ = oracle.pagerank(np.ones(n, np.uint32), ptr, in_cols, 0.85, 1e-14, 1000, threads=16)
    assert np.abs(r - 1 / n).max() < 1e-18
    d, m = 0.85, 5000
    nn = m + 1
    od = np.array([m] + [1] * m, np.uint32)
    ptr = np.array([0] + [m + i for i in range(m)], np.uint64)
    in_cols = np.array(list(range(1, m + 1)) + [0] * m, np.uint32)
    r, _, _ = oracle.pagerank(od, ptr, in_cols, d, 1e-16, 10000, threads=16)
    leaf = (1 + d / m) / (nn * (1 + d))
    center = (1 - d) / nn + d * m * leaf
    assert abs(r[0] - center) < 1e-13 and np.abs(r[1:] - leaf).max() < 1e-16


def test_sssp_saturates_below_unreached():
    """distance + 1 is capped at 0xFFFFFFFE (0xFFFFFFFF stays 'unreached'):
    a chain whose sums cross 2^32, closed forms computed here."""
    ws = [0x7FFFFFFF, 0x7FFFFFFF, 0x7FFFFFFF, 5, 0xFFFFFFF0]
    n = len(ws) + 2
    edges = [(i, i + 1) for i in range(len(ws))]
    ptr, dst, w = csr_from_edges(n, edges, ws)
    got = oracle.sssp(ptr, dst, w, 0)
    dist, acc = [0], 0
    for x in ws:
        acc += x
        dist.append(acc)
    want = [min(d + 1, 0xFFFFFFFE) for d in dist] + [U]
    assert got.tolist() == want
    # the largest distance that is not capped: d + 1 = 0xFFFFFFFD
    ptr, dst, w = csr_from_edges(3, [(0, 1), (1, 2)], [0xFFFFFFFC, 1])
    assert oracle.sssp(ptr, dst, w, 0).tolist() == [1, 0xFFFFFFFD, 0xFFFFFFFE]


def test_loop_seconds_reported():
    ptr, cols, _ = csr_from_edges(4, [(0, 1), (1, 2), (2, 3)])
    oracle.bfs(ptr, cols, 0)
    t = oracle.loop_seconds()
    assert 0.0 <= t < 1.0
