"""bench.py's contract on the host: one GPU only (--gpus above 1 exits with a
message and no JSON line), every configuration line it names exists, and the
HBM traffic it reports for the headline roofline comes from the committed
rocprofv3 counter summaries with their calibration factors, and the gather's
texture-path rate is bytes per clock per CU."""
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


def test_texture_path_rate():
    # 5.12 M samples x 32 coils x 125 taps in 5.585 ms: 29.3 TB/s of tap reads,
    # 47.7 bytes per clock on each of 256 CUs at 2.4 GHz
    t = bench.texture_path(5.12e6 * 32 * 125 / 5.585e-3)
    assert abs(t["achieved"] - 47.74) < 0.01 and t["peak"] == 64.0
    assert abs(t["frac"] - t["achieved"] / 64.0) < 1e-12


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
