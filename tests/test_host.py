"""Host-side checks that need no GPU: generators against an independent numpy
restatement of the published algorithms, reference CSR byte format, the
static-region planner, error statuses, and the C ABI export list."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import onegraph_amd as og

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
GOLD = np.uint64(0x9E3779B97F4A7C15)


# ----------------------------------------------------------- restatement
def mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def key_of(seed):
    return mix64(np.uint64(seed) ^ np.uint64(0x6A09E667F3BCC909))


def draw(key, ctr):
    with np.errstate(over="ignore"):
        return mix64(key + (np.asarray(ctr, np.uint64) + np.uint64(1)) * GOLD)


def to_csr(V, s, d):
    keep = s != d
    k = np.unique((s[keep].astype(np.uint64) << np.uint64(32)) | d[keep].astype(np.uint64))
    src = (k >> np.uint64(32)).astype(np.int64)
    dst = (k & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(src, minlength=V))[:-1]]).astype(np.uint64)
    return ptr, dst


def rmat_numpy(scale, ef, a, b, c, seed):
    """R-MAT (Chakrabarti et al. 2004), Graph500 quadrant order, MSB first,
    one 32-bit uniform per level from a SplitMix64 counter stream."""
    M = ef << scale
    key = key_of(seed)
    W = (scale + 1) // 2
    i = np.arange(M, dtype=np.uint64)
    TA = np.uint64(int(np.floor(a * 4294967296.0)))
    TAB = np.uint64(int(np.floor((a + b) * 4294967296.0)))
    TABC = np.uint64(int(np.floor((a + b + c) * 4294967296.0)))
    s = np.zeros(M, np.uint64)
    d = np.zeros(M, np.uint64)
    for l in range(scale):
        r = draw(key, i * np.uint64(W) + np.uint64(l // 2))
        x = (r >> np.uint64(32 * (l & 1))) & np.uint64(0xFFFFFFFF)
        sb = (x >= TAB).astype(np.uint64)
        db = (((x >= TA) & (x < TAB)) | (x >= TABC)).astype(np.uint64)
        s = (s << np.uint64(1)) | sb
        d = (d << np.uint64(1)) | db
    return to_csr(1 << scale, s, d)


def weights_numpy(E, seed):
    r = draw(key_of(seed), np.arange(E, dtype=np.uint64)) >> np.uint64(32)
    return (np.uint64(1) + ((r * np.uint64(63)) >> np.uint64(32))).astype(np.uint32)


def er_numpy(n, avg, seed):
    M = n * avg // 2
    key = key_of(seed)
    j = np.arange(M, dtype=np.uint64)

    def uni(r):  # floor(r * n / 2^64) via 32-bit halves
        hi, lo = r >> np.uint64(32), r & np.uint64(0xFFFFFFFF)
        return (hi * np.uint64(n) + ((lo * np.uint64(n)) >> np.uint64(32))) >> np.uint64(32)

    u = uni(draw(key, np.uint64(2) * j))
    v = uni(draw(key, np.uint64(2) * j + np.uint64(1)))
    return to_csr(n, np.concatenate([u, v]), np.concatenate([v, u]))


# ----------------------------------------------------------- generators
@pytest.mark.parametrize("scale,ef,seed", [(6, 4, 1), (10, 16, 1002), (11, 8, 7)])
def test_rmat_matches_restatement(scale, ef, seed):
    g = og.Graph.rmat(scale, ef, 0.57, 0.19, 0.19, seed=seed)
    ptr, cols, _, _ = g.export()
    rp, rc = rmat_numpy(scale, ef, 0.57, 0.19, 0.19, seed)
    assert np.array_equal(ptr, rp) and np.array_equal(cols, rc)


def test_weights_and_er_match_restatement():
    g = og.Graph.rmat(9, 8, seed=3, weight_seed=2002)
    ptr, to, w, _ = g.export()
    assert w.min() >= 1 and w.max() <= 63
    assert np.array_equal(w, weights_numpy(len(to), 2002))
    e = og.Graph.er(3000, 10, seed=4)
    ptr, cols, _, _ = e.export()
    rp, rc = er_numpy(3000, 10, 4)
    assert np.array_equal(ptr, rp) and np.array_equal(cols, rc)


def test_generator_invariants():
    g = og.Graph.rmat(12, 16, seed=1)
    ptr, cols, _, _ = g.export()
    V, E = len(ptr), len(cols)
    src = np.repeat(np.arange(V), np.diff(np.append(ptr.astype(np.int64), E)))
    assert np.all(src != cols)  # no self-loops
    key = src.astype(np.int64) * V + cols
    assert np.all(np.diff(key) > 0)  # rows sorted, no duplicates
    deg = np.bincount(src, minlength=V)
    assert deg.argmax() == 0  # a = .57 concentrates edges on low ids
    assert g.info()["max_degree"] == deg.max()
    e = og.Graph.er(2048, 16, seed=2)
    ptr, cols, _, _ = e.export()
    s = np.repeat(np.arange(2048), np.diff(np.append(ptr.astype(np.int64), len(cols))))
    fwd = set(zip(s.tolist(), cols.tolist()))
    assert all((b, a) in fwd for a, b in fwd)  # symmetric
    assert abs(len(cols) / 2048 - 16) < 0.5


def test_generator_determinism():
    a = og.Graph.rmat(12, 8, seed=5).export()
    b = og.Graph.rmat(12, 8, seed=5).export()
    c = og.Graph.rmat(12, 8, seed=6).export()
    assert np.array_equal(a[1], b[1]) and not np.array_equal(a[1], c[1])


def test_to_pull():
    g = og.Graph.rmat(9, 8, seed=9)
    ptr, cols, _, _ = g.export()
    p = g.to_pull()
    pptr, pin, _, od = p.export()
    V, E = len(ptr), len(cols)
    src = np.repeat(np.arange(V), np.diff(np.append(ptr.astype(np.int64), E)))
    assert np.array_equal(od, np.bincount(src, minlength=V))
    dst = np.repeat(np.arange(V), np.diff(np.append(pptr.astype(np.int64), E)))
    assert sorted(zip(src.tolist(), cols.tolist())) == sorted(zip(pin.tolist(), dst.tolist()))


def test_find_big_node():
    g = og.Graph.rmat(14, 16, seed=1)
    ptr, cols, _, _ = g.export()
    deg = np.diff(np.append(ptr.astype(np.int64), len(cols)))
    big = [i for i in range(1, len(ptr) - 1) if ptr[i + 1] - ptr[i] > 3000]
    assert g.find_big_node(0) == (big[0] if big else 0)
    if len(big) > 1:
        assert g.find_big_node(big[0]) == big[1]
    assert g.find_big_node(len(ptr) - 2) == len(ptr) - 2  # none left: returns `last`


# ----------------------------------------------------------- file format
def test_file_round_trip(tmp_path):
    g = og.Graph.rmat(8, 8, seed=2, weight_seed=3)
    ptr, to, w, _ = g.export()
    f = str(tmp_path / "w.bwcsr")
    g.save(f)
    raw = open(f, "rb").read()
    V, E = struct.unpack("<QQ", raw[:16])
    assert (V, E) == (len(ptr), len(to))
    assert len(raw) == 16 + 8 * V + 8 * E
    assert np.array_equal(np.frombuffer(raw[16:16 + 8 * V], np.uint64), ptr)
    aos = np.frombuffer(raw[16 + 8 * V:], np.uint32).reshape(-1, 2)
    assert np.array_equal(aos[:, 0], to) and np.array_equal(aos[:, 1], w)
    h = og.Graph.load(f, og.KIND_WEIGHTED)
    p2, t2, w2, _ = h.export()
    assert np.array_equal(p2, ptr) and np.array_equal(t2, to) and np.array_equal(w2, w)
    # plain and pull layouts
    pg = og.Graph.rmat(8, 8, seed=2)
    pf = str(tmp_path / "p.bcsr")
    pg.save(pf)
    assert np.array_equal(og.Graph.load(pf).export()[1], pg.export()[1])
    pl = pg.to_pull()
    lf = str(tmp_path / "p.prcsr")
    pl.save(lf)
    raw = open(lf, "rb").read()
    V, E = struct.unpack("<QQ", raw[:16])
    assert np.array_equal(np.frombuffer(raw[16:16 + 4 * V], np.uint32), pl.export()[3])
    back = og.Graph.load(lf, og.KIND_PULL).export()
    assert np.array_equal(back[1], pl.export()[1]) and np.array_equal(back[3], pl.export()[3])


def _write(path, V, E, ptr, payload):
    with open(path, "wb") as f:
        f.write(struct.pack("<QQ", V, E))
        f.write(np.asarray(ptr, np.uint64).tobytes())
        f.write(np.asarray(payload, np.uint32).tobytes())


@pytest.mark.parametrize("case,status", [
    ("missing", "OG_EIO"), ("short_header", "OG_EIO"), ("truncated", "OG_EIO"), ("trailing", "OG_EFORMAT"),
    ("bad_col", "OG_EFORMAT"), ("decreasing", "OG_EFORMAT"), ("zero_v", "OG_EFORMAT"), ("overrun", "OG_EFORMAT"),
])
def test_load_errors(tmp_path, case, status):
    f = str(tmp_path / "bad.bcsr")
    if case == "short_header":
        open(f, "wb").write(b"\x01\x00")
    elif case == "truncated":
        _write(f, 3, 4, [0, 1, 2], [1, 2, 0])
    elif case == "trailing":
        _write(f, 3, 2, [0, 1, 2], [1, 2, 0])
    elif case == "bad_col":
        _write(f, 3, 3, [0, 1, 2], [1, 7, 0])
    elif case == "decreasing":
        _write(f, 3, 3, [0, 2, 1], [1, 2, 0])
    elif case == "zero_v":
        _write(f, 0, 0, [], [])
    elif case == "overrun":
        _write(f, 3, 3, [0, 1, 4], [1, 2, 0])
    with pytest.raises(og.OgError) as e:
        og.Graph.load(f if case != "missing" else str(tmp_path / "nope"))
    assert e.value.name == status


def test_from_csr_edge_cases():
    g = og.Graph.from_csr(np.zeros(5, np.uint64), np.zeros(0, np.uint32))
    assert g.info() == {"num_vertices": 5, "num_edges": 0, "kind": og.KIND_PLAIN, "max_degree": 0}
    with pytest.raises(og.OgError) as e:
        og.Graph.from_csr(np.array([0, 1], np.uint64), np.array([0, 9], np.uint32))
    assert e.value.name == "OG_EFORMAT"


# ----------------------------------------------------------- planner
def test_planner_hand_cases():
    # V=4, rows of 3,2,4,1 edges; budget 10 edges of 4 bytes
    ptr = np.array([0, 3, 5, 9], np.uint64)
    p = og.plan_compute(ptr, 10, 4, edge_budget_bytes=40, advisek=0.5)
    assert p["total_budget_edges"] == 10 and p["static_edges"] == 5 and p["staging_edges"] == 5
    assert p["static_vertices"] == 2 and p["overload_edges"] == 5 and p["vsplit"] == 2
    # auto rule: (10 - E/total)/9 = 1 -> whole graph static
    p = og.plan_compute(ptr, 10, 4, edge_budget_bytes=40)
    assert abs(p["advisek_used"] - 1.0) < 1e-7 and p["static_edges"] == 10 and p["vsplit"] == 4
    # budget 2x smaller than E: (10 - 2)/9 = 0.888..
    p = og.plan_compute(ptr, 10, 4, edge_budget_bytes=20)
    assert abs(p["advisek_used"] - 8 / 9) < 1e-6 and p["static_edges"] == int(float(np.float32(8 / 9)) * 5)
    # advisek = 1 with a budget of exactly E edges keeps every edge static even
    # when E is not representable in float (E = 2^24 + 3 here)
    big = np.array([0, (1 << 24) + 1], np.uint64)
    p = og.plan_compute(big, (1 << 24) + 3, 4, edge_budget_bytes=((1 << 24) + 3) * 4, advisek=1.0)
    assert p["static_edges"] == (1 << 24) + 3 and p["vsplit"] == 2 and p["staging_edges"] == 0
    # E/total = 10: the rule gives exactly 0
    p = og.plan_compute(ptr, 10, 4, edge_budget_bytes=4)
    assert p["advisek_used"] == 0.0 and p["static_edges"] == 0
    # E/total > 10: the rule goes negative and the reference substitutes 0.5
    p30 = np.array([0, 10, 20, 25], np.uint64)
    p = og.plan_compute(p30, 30, 4, edge_budget_bytes=8)
    assert p["advisek_used"] == 0.5 and p["static_edges"] == 1 and p["vsplit"] == 0
    # weighted edges take 8 bytes
    p = og.plan_compute(ptr, 10, 8, edge_budget_bytes=40, advisek=1.0)
    assert p["total_budget_edges"] == 5 and p["static_edges"] == 5 and p["vsplit"] == 2
    # reference default: 12 GiB - 17*4*V bytes
    p = og.plan_compute(ptr, 10, 4)
    assert p["total_budget_edges"] == ((12 << 30) - 17 * 4 * 4) // 4
    # zero-degree vertex at offset 0: the reference's ptr+deg-1 wraps and
    # leaves it out of the static count (GraphMeta.h:346)
    p = og.plan_compute(np.array([0, 0, 2], np.uint64), 4, 4, edge_budget_bytes=16, advisek=1.0)
    assert p["static_vertices"] == 2 and p["vsplit"] == 3


def test_planner_brute_force():
    rng = np.random.default_rng(1)
    for _ in range(200):
        V = int(rng.integers(1, 30))
        deg = rng.integers(0, 6, V)
        E = int(deg.sum())
        ptr = np.concatenate([[0], np.cumsum(deg)[:-1]]).astype(np.uint64)
        budget = int(rng.integers(1, 3 * E + 5)) * 4
        k = float(rng.choice([0.0, 0.3, 0.8, 1.0]))
        p = og.plan_compute(ptr, E, 4, edge_budget_bytes=budget, advisek=k)
        S = p["static_edges"]
        ends = np.append(ptr[1:], E)
        # u64 arithmetic as in the reference: ptr + deg - 1 wraps for an empty row at 0
        inside = [(int(ptr[i]) < S) and ((int(ptr[i]) + int(deg[i]) - 1) % (1 << 64) < S) for i in range(V)]
        assert p["static_vertices"] == sum(inside)
        assert p["overload_edges"] == E - int(sum(deg[i] for i in range(V) if inside[i]))
        assert p["staging_edges"] == p["total_budget_edges"] - S
        vs = p["vsplit"]
        assert all(ends[:vs] <= S) and (vs == V or ends[vs] > S)


# ----------------------------------------------------------- ABI
def test_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "onegraph.h")).read()
    names = set(re.findall(r"\b(og_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 15
    lib = og.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", og.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (og_[a-z0-9_]+)", out))
    assert names <= exported


def test_no_device_errors_are_statuses():
    if og.device_count() > 0:
        pytest.skip("GPU present")
    g = og.Graph.rmat(6, 4, seed=1)
    with pytest.raises(og.OgError) as e:
        g.prepare("ascetic")
    assert e.value.name == "OG_EHIP"
    with pytest.raises(og.OgError) as e:
        g.bfs(0)
    assert e.value.name == "OG_ESTATE"


def test_levels_checksum_matches_restatement():
    rng = np.random.default_rng(2)
    x = rng.integers(1, 40, 5000).astype(np.uint32)
    x[rng.random(5000) < 0.3] = og.UNREACHED
    keep = x != og.UNREACHED
    v = np.arange(5000, dtype=np.uint64)[keep]
    with np.errstate(over="ignore"):
        ref = int(mix64((v << np.uint64(32)) | x[keep].astype(np.uint64)).sum(dtype=np.uint64))
    assert og.levels_checksum(x) == ref


def test_plan_flags_match_header():
    """The Python PLAN_* constants are the header's OG_PLAN_* values."""
    hdr = open(os.path.join(ROOT, "include", "onegraph.h")).read()
    defs = dict(re.findall(r"#define OG_PLAN_([A-Z_]+) (\d+)u", hdr))
    assert set(defs) >= {"NO_OVERLAP", "ZERO_COPY", "NO_PACK", "NO_GRAPH", "DIRECTION_OPT", "NO_FUSE"}
    for name, value in defs.items():
        assert getattr(og, "PLAN_" + name) == int(value), name
    assert len(set(defs.values())) == len(defs)  # distinct bits
    sdefs = dict(re.findall(r"#define OG_SERVER_([A-Z_]+) (0x[0-9a-f]+)u", hdr))
    assert set(sdefs) == {"NO_SSSP_LANES", "SSSP_BATCH", "TEST_FAIL_SSSP"}
    for name, value in sdefs.items():
        assert getattr(og, "SERVER_" + name) == int(value, 16), name
