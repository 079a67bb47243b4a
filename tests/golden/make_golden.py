"""Writes the golden fixtures in tests/golden/*.npz.

The expected outputs come from independent library routines in this
container (scipy.sparse.csgraph, networkx), never from this repository's
code: inputs are small seeded random graphs built with numpy.

  bfs_*.npz   levels = hop distance + 1 (scipy shortest_path, unweighted),
              unreached 0xFFFFFFFF
  sssp_*.npz  distances + 1 (scipy dijkstra on integer weights, exact in FP64)
  cc_*.npz    undirected: min vertex id of the scipy connected component;
              directed: min id over all vertices that reach v (scipy
              reachability), the fixed point of min-label push
  pr_*.npz    networkx.pagerank(alpha=0.85, tol=1e-15), dangling mass spread
              uniformly

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import networkx as nx
import numpy as np
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components, dijkstra, shortest_path

OUT = os.path.dirname(os.path.abspath(__file__))
U = 0xFFFFFFFF


def random_graph(rng, V, E, directed=True, self_loops=False):
    src = rng.integers(0, V, E)
    dst = rng.integers(0, V, E)
    if not self_loops:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    if not directed:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
    key = np.unique(src.astype(np.uint64) << np.uint64(32) | dst.astype(np.uint64))
    s = (key >> np.uint64(32)).astype(np.int64)
    d = (key & np.uint64(0xFFFFFFFF)).astype(np.int64)
    counts = np.bincount(s, minlength=V)
    ptr = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint64)  # V entries
    return ptr, d.astype(np.uint32), s, d


def levels_from(dist):
    out = np.full(len(dist), U, np.uint32)
    fin = np.isfinite(dist)
    out[fin] = dist[fin].astype(np.uint64) + 1
    return out


def main():
    rng = np.random.default_rng(20241202)
    cases = [("sparse", 300, 600, True), ("dense", 120, 2400, True), ("undirected", 400, 500, False),
             ("chain_heavy", 1000, 1100, True), ("tiny", 5, 4, True)]
    for name, V, E, directed in cases:
        ptr, cols, s, d = random_graph(rng, V, E, directed)
        Ecnt = len(cols)
        A = csr_matrix((np.ones(Ecnt), (s, d)), shape=(V, V))
        w = rng.integers(1, 64, Ecnt).astype(np.uint32)
        W = csr_matrix((w.astype(np.float64), (s, d)), shape=(V, V))
        deg = np.bincount(s, minlength=V)
        sources = [int(np.flatnonzero(deg > 0)[0])] + [int(x) for x in rng.choice(V, 3, replace=False)]
        bfs = np.stack([levels_from(shortest_path(A, method="D", unweighted=True, indices=x)) for x in sources])
        sssp = np.stack([levels_from(dijkstra(W, indices=x)) for x in sources])
        if directed:
            reach = np.isfinite(shortest_path(A, method="D", unweighted=True))  # reach[u, v]
            cc = np.array([np.flatnonzero(reach[:, v]).min() for v in range(V)], np.uint32)
        else:
            _, comp = connected_components(A, directed=False)
            mins = np.full(comp.max() + 1, V, np.int64)
            np.minimum.at(mins, comp, np.arange(V))
            cc = mins[comp].astype(np.uint32)
        G = nx.DiGraph()
        G.add_nodes_from(range(V))
        G.add_edges_from(zip(s.tolist(), d.tolist()))
        pr = nx.pagerank(G, alpha=0.85, tol=1e-15, max_iter=10000)
        pr = np.array([pr[i] for i in range(V)], np.float64)
        np.savez(os.path.join(OUT, f"graph_{name}.npz"), ptr=ptr, cols=cols, weights=w, sources=np.array(sources),
                 bfs=bfs, sssp=sssp, cc=cc, pagerank=pr, directed=np.array(directed))
        print(name, V, Ecnt, "reached from s0:", int((bfs[0] != U).sum()))


if __name__ == "__main__":
    main()
