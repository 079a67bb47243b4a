"""The drop-in command line: reference flag spellings, errors as exit codes,
and (on the GPU) result dumps equal to the oracle."""
import subprocess

import numpy as np
import pytest

import onegraph_amd as og
import oracle


def run(*args, timeout=120):
    return subprocess.run([og.CLI_PATH, *args], capture_output=True, text=True, timeout=timeout)


def test_usage_and_errors(tmp_path):
    r = run("--help")
    assert r.returncode == 0 and "--type" in r.stdout
    assert run("--bogus").returncode == 2
    assert run("--type", "bfs").returncode == 2          # no -f
    assert run("--type", "bfs", "-f").returncode == 2    # flag without value
    assert run("--type", "xx", "-f", "x").returncode == 2
    assert run("--type", "bfs", "-f", "x", "--loader", "fast").returncode == 2
    r = run("--type", "bfs", "-f", str(tmp_path / "missing.bcsr"))
    assert r.returncode == 1 and "OG_EIO" not in r.stdout  # message goes to stderr
    assert "cannot open" in r.stderr


def test_generate_and_save(tmp_path):
    f = str(tmp_path / "g.bwcsr")
    r = run("-type", "sssp", "--generate", "rmat:9:8:3:4", "--save", f)
    assert r.returncode == 0, r.stderr
    g = og.Graph.load(f, og.KIND_WEIGHTED)
    ref = og.Graph.rmat(9, 8, seed=3, weight_seed=4)
    assert np.array_equal(g.export()[1], ref.export()[1]) and np.array_equal(g.export()[2], ref.export()[2])


@pytest.mark.gpu
@pytest.mark.parametrize("algo,flags", [
    ("bfs", ["--advisek", "0.8", "--budget-fraction", "0.25"]),
    ("sssp", ["-adviseK", "0.8", "--budget-fraction", "0.25"]),
    ("cc", ["--adviseK", "0.5", "--budget-fraction", "0.1", "--approach", "subway"]),
    ("bfs", ["--approach", "usm", "--method", "1"]),
    ("sssp", ["--advisek", "0.8", "--budget-fraction", "0.25", "--loader", "zerocopy", "--delta", "16",
              "--split-weight", "16", "--autotune"]),
    ("bfs", ["--advisek", "1.0", "--budget-fraction", "1.0", "--direction-opt", "--repeat", "2"]),
    ("cc", ["--advisek", "0.8", "--budget-fraction", "0.25", "--loader", "zerocopy"]),
])
def test_cli_runs_match_oracle(gpu, tmp_path, algo, flags):
    f = str(tmp_path / "g.csr")
    kind = og.KIND_WEIGHTED if algo == "sssp" else og.KIND_PLAIN
    g = og.Graph.rmat(12, 8, seed=5, weight_seed=6 if kind == og.KIND_WEIGHTED else 0)
    g.save(f)
    ptr, to, w, _ = g.export()
    out = str(tmp_path / "out.bin")
    src = int(np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), len(to))) > 0)[3])
    r = run("--type", algo, "-f", f, "-r", str(src), "--out", out, "--json", *flags)
    assert r.returncode == 0, r.stderr
    assert "totalProcess time is" in r.stdout and "nodeSum :" in r.stdout
    got = np.fromfile(out, np.uint32)
    if algo == "bfs":
        ref = oracle.bfs(ptr, to, src)[0]
    elif algo == "sssp":
        ref = oracle.sssp(ptr, to, w, src)
    else:
        ref = oracle.cc(ptr, to)[0]
    assert np.array_equal(got, ref)


@pytest.mark.gpu
def test_cli_bignode_and_pagerank(gpu, tmp_path):
    f = str(tmp_path / "g.bcsr")
    g = og.Graph.rmat(14, 16, seed=9)
    g.save(f)
    r = run("--type", "bfs", "-f", f, "--bignode-runs", "3")
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("testIndex") == 3
    pf = str(tmp_path / "g.prcsr")
    p = g.to_pull()
    p.save(pf)
    out = str(tmp_path / "pr.bin")
    r = run("--type", "pr", "-f", pf, "--advisek", "0.8", "--budget-fraction", "0.25", "--fp64", "--out", out)
    assert r.returncode == 0, r.stderr
    pptr, pin, _, od = p.export()
    ref, _, _ = oracle.pagerank(od, pptr, pin)
    assert np.abs(np.fromfile(out, np.float64) - ref).max() < 1e-9


@pytest.mark.gpu
def test_cli_text_input_and_check(gpu, tmp_path):
    """A text edge list runs directly (--text) and --check compares the result
    with a dump: the oracle's levels pass, a planted difference exits 3."""
    rng = np.random.default_rng(17)
    V, M = 3000, 40000
    src, dst = rng.integers(0, V, M), rng.integers(0, V, M)
    t = tmp_path / "e.txt"
    np.savetxt(t, np.stack([src, dst], 1), fmt="%d")
    g = og.Graph.from_edges(src, dst, num_vertices=V)
    ptr, to, _, _ = g.export()
    s = int(src[0])
    ref = oracle.bfs(ptr, to, s)[0]
    good, bad = tmp_path / "good.u32", tmp_path / "bad.u32"
    ref.tofile(good)
    r = run("--type", "bfs", "-f", str(t), "--text", "-r", str(s), "--budget-fraction", "0.3",
            "--advisek", "0.8", "--check", str(good))
    assert r.returncode == 0, r.stderr
    assert "values match" in r.stdout
    wrong = ref.copy()
    wrong[int(np.flatnonzero(ref != og.UNREACHED)[-1])] += 1
    wrong.tofile(bad)
    r = run("--type", "bfs", "-f", str(t), "--text", "-r", str(s), "--check", str(bad))
    assert r.returncode == 3 and "values differ" in r.stderr
    g.close()
