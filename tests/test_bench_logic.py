"""Host-side logic of bench.py's roofline fields (no GPU): the dominant
kernel class, the host-link figure of zero-copy runs, and the fused round
kernel (static bins timed inside the dynamic class)."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

BASE = {"relax_static_launches": 10, "relax_static_bytes": 1e9, "relax_dynamic_kernel_ms": 10.0,
        "relax_dynamic_launches": 10, "relax_dynamic_bytes": 2e9, "frontier_kernel_ms": 1.0,
        "frontier_launches": 10, "frontier_bytes": 1e8, "copy_ms": 0.0, "h2d_bytes": 4e8}


@pytest.mark.parametrize("static_ms", [0.0, 5.0])
def test_zero_copy_roofline(static_ms):
    st = dict(BASE, relax_static_kernel_ms=static_ms)
    roof, link, kt = bench.roofline_of([st, st], zero_copy=True, traffic=123)
    fused = static_ms == 0.0
    # two runs: 8e8 link bytes over 20 ms of the dynamic class = 40 GB/s
    assert roof["bound"] == "host-link" and roof["achieved"] == pytest.approx(40.0)
    assert roof["frac"] == pytest.approx(40.0 / bench.LINK_PEAK_GBS, abs=1e-4)
    assert roof["link_bytes_per_launch"] == pytest.approx(4e7)
    assert roof["avg_launch_us"] == pytest.approx(1000.0)
    assert roof["traffic"] == 123
    # the fused launch carries both regions' algorithmic bytes
    hbm_bytes = (2e9 + (1e9 if fused else 0.0)) / 10
    assert roof["hbm"]["bytes_per_launch"] == pytest.approx(hbm_bytes)
    assert ("k_relax_round" in roof["kernel"]) == fused
    assert link["measured_on"].startswith("k_relax_round" if fused else "k_relax_dyn")
    assert ("relax_static" in kt) == (not fused)


def test_in_hbm_roofline():
    st = dict(BASE, relax_static_kernel_ms=20.0, relax_dynamic_kernel_ms=0.0, relax_dynamic_launches=0,
              h2d_bytes=0.0)
    roof, link, kt = bench.roofline_of([st], zero_copy=False)
    assert roof["bound"] == "hbm" and "k_relax_group" in roof["kernel"]
    assert roof["achieved"] == pytest.approx(1e9 / 20e-3 / 1e9)
    assert link is None
