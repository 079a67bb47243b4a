"""Outside callers of the library: a plain C program on the C ABI and a C++
program on the host layer (the reference's driver names), compiled with
gcc / g++ against include/ and linked against libonegraph.so, as a
maintainer swapping the reference's programs would build them."""
import os
import subprocess

import numpy as np
import pytest

import onegraph_amd as og
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.dirname(og.LIB_PATH)


def _build(tmp_path, src, exe, compiler, std):
    og.lib()  # the library is built
    out = str(tmp_path / exe)
    cmd = [compiler, std, "-O1", "-Wall", "-o", out, os.path.join(HERE, "integration", src),
           "-I" + os.path.join(ROOT, "include"), "-L" + LIBDIR, "-lonegraph", "-Wl,-rpath," + LIBDIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.fixture(scope="module")
def callers(tmp_path_factory):
    d = tmp_path_factory.mktemp("callers")
    return (_build(d, "caller.c", "caller", "gcc", "-std=c11"),
            _build(d, "caller.cpp", "caller_cpp", "g++", "-std=c++17"))


@pytest.fixture(scope="module")
def wgraph(tmp_path_factory):
    f = str(tmp_path_factory.mktemp("g") / "g.wbcsr")
    g = og.Graph.rmat(11, 8, seed=21, weight_seed=22)
    g.save(f)
    return f, g


def test_c_caller_plans_like_the_library(callers, wgraph):
    c, _ = callers
    f, g = wgraph
    ptr, to, _, _ = g.export()
    for budget, k in ((len(to) * 8 // 4, 0.8), (len(to) * 8, 1.0), (len(to) * 2, 0.5)):
        r = subprocess.run([c, "plan", f, str(budget), str(k)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        p = og.plan_compute(ptr, len(to), 8, edge_budget_bytes=budget, advisek=k)
        want = f"total {p['total_budget_edges']} static {p['static_edges']} staging {p['staging_edges']} " \
               f"vsplit {p['vsplit']}"
        assert r.stdout.strip() == want


def test_callers_report_errors_as_statuses(callers, tmp_path):
    c, cpp = callers
    r = subprocess.run([c, "plan", str(tmp_path / "missing"), "1", "1"], capture_output=True, text=True)
    assert r.returncode == 2 and "status 2" in r.stderr  # OG_EIO
    r = subprocess.run([cpp, str(tmp_path / "missing"), "0", "0.8", str(tmp_path / "o")], capture_output=True,
                       text=True)
    assert r.returncode == 2  # OG_EIO from bfs_opt, no exit() inside the library


def test_cpp_caller_links_the_drivers(callers, tmp_path):
    """bfs_opt comes from libonegraph.so; on a host without a GPU it returns
    OG_EHIP instead of exiting."""
    if og.device_count() > 0:
        pytest.skip("GPU present: test_cpp_caller_bfs_matches_oracle runs it")
    _, cpp = callers
    f = str(tmp_path / "g.bcsr")
    og.Graph.rmat(10, 8, seed=24).save(f)
    r = subprocess.run([cpp, f, "0", "0.8", str(tmp_path / "o")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 7, r.stderr  # OG_EHIP
    assert "no HIP device" in r.stderr


@pytest.mark.gpu
def test_cpp_caller_bfs_matches_oracle(gpu, callers, tmp_path):
    _, cpp = callers
    g = og.Graph.rmat(12, 8, seed=23)
    f = str(tmp_path / "g.bcsr")
    g.save(f)
    ptr, cols, _, _ = g.export()
    src = int(np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), len(cols))) > 0)[5])
    out = str(tmp_path / "lv.bin")
    r = subprocess.run([cpp, f, str(src), "0.8", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "totalProcess time is" in r.stdout
    ref, _, _ = oracle.bfs(ptr, cols, src)
    assert np.array_equal(np.fromfile(out, np.uint32), ref)


@pytest.mark.gpu
def test_c_caller_sssp_matches_oracle(gpu, callers, wgraph, tmp_path):
    c, _ = callers
    f, g = wgraph
    ptr, to, w, _ = g.export()
    src = int(np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), len(to))) > 0)[2])
    out = str(tmp_path / "d.bin")
    r = subprocess.run([c, "sssp", f, str(src), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(np.fromfile(out, np.uint32), oracle.sssp(ptr, to, w, src))


@pytest.mark.gpu
def test_c_caller_serves_like_the_library(gpu, callers, wgraph):
    """The server through the C ABI from a C program: its SSSP (on a lane)
    and BFS answers carry the checksums of og_sssp / og_bfs from the same
    source."""
    c, _ = callers
    f, g = wgraph
    ptr, to, _, _ = g.export()
    src = int(np.flatnonzero(np.diff(np.append(ptr.astype(np.int64), len(to))) > 0)[3])
    r = subprocess.run([c, "serve", f, str(src), "-"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok sssp_slots 2"), r.stdout
