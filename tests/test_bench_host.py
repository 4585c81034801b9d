"""Host-side pieces of bench.py (no GPU): the lookup of PMC-measured HBM
traffic per kernel launch."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_measured_traffic_matches_template_suffix(tmp_path):
    p = tmp_path / "traffic.json"
    p.write_text(json.dumps({"kernels": {
        "spray3_level<2, 2>": {"fetch_bytes": 100, "write_bytes": 20, "dispatches": 3},
        "spray3_level<2, 1>": {"fetch_bytes": 50, "write_bytes": 10, "dispatches": 3},
        "smooth_contig<1, 0>": {"fetch_bytes": 7, "write_bytes": 1, "dispatches": 3},
        "cg_start_kernel": {"fetch_bytes": 4, "write_bytes": 4, "dispatches": 3}}}))
    assert bench.measured_traffic("spray3_level<2>", str(p)) == 120
    assert bench.measured_traffic("spray3_level<1>", str(p)) == 60
    assert bench.measured_traffic("spray3_level", str(p)) == 90
    assert bench.measured_traffic("smooth_contig<1,0>", str(p)) == 8
    assert bench.measured_traffic("cg_start_kernel", str(p)) == 8
    assert bench.measured_traffic("no_such_kernel", str(p)) is None
    assert bench.measured_traffic("x", str(tmp_path / "absent.json")) is None


def test_committed_traffic_file_parses():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(path):
        # the library's timer names for the warm path's level launches
        # (sources, prefetch distance) must resolve in the committed file
        for name in ("spray3_level<2,6>", "spray3_level<1,8>"):
            t = bench.measured_traffic(name, path)
            assert t is not None and t > 0, name
