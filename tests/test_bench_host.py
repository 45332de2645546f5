"""bench.py's contract on the host: one GPU only (--gpus above 1 exits with a
message and no JSON line), every configuration line it names exists, and the
HBM traffic it reports for the headline roofline comes from the committed
rocprofv3 counter summaries with their calibration factors."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_config_lines_name_functions():
    assert len(bench.CONFIG_LINES) == 4
    for name, fn in bench.CONFIG_LINES.items():
        assert callable(getattr(bench, fn, None)), name


def test_profiled_traffic_from_committed_summaries():
    t = bench.profiled_traffic(bench.KERNEL_NAMES["gather"])
    assert t is not None
    cal = t["calibration"]
    assert cal["FETCH_SIZE"] > 0 and cal["WRITE_SIZE"] > 0
    want = 1024.0 * (t["fetch_kib"] * cal["FETCH_SIZE"] + t["write_kib"] * cal["WRITE_SIZE"])
    assert abs(t["bytes"] - want) <= 1e-9 * want
    for f in t["files"]:
        assert os.path.exists(os.path.join(ROOT, f))
    # 256^3 x 32 coils: the gather's algorithmic bytes are 9.78 GB; the
    # counters hold its cells re-read across XCDs, within a small factor
    assert 1.0 < t["bytes"] / 9.781248e9 < 10.0
