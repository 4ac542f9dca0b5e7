"""bench.py's roofline arithmetic on the CPU: the blocked kernel's compulsory
bytes per launch, the per-kind table built from HIP-event timings, and the
refusal of more than one GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_pipe_bytes_per_launch():
    # two fields in and out plus the model; T virtual sources; the gradient read and written once
    assert bench.pipe_bytes(3) == {"fwd": 20, "fwd_virt": 36, "recomp": 36, "adj": 40}
    assert bench.pipe_bytes(2)["adj"] == 36 and bench.pipe_bytes(4)["recomp"] == 40


def test_kernel_table_rates_and_shares():
    cells = 194_052_096
    # 2667 launches of 3 steps: timing() counts steps, milliseconds summed
    tm = {"fwd": (2667 * 1.0, 8001), "fwd_virt": (0.0, 0), "recomp": (2667 * 1.5, 8001), "adj": (2667 * 2.0, 8001)}
    per = bench.kernel_table(tm, cells, 3)
    assert set(per) == {"fwd", "recomp", "adj"}
    assert per["adj"]["avg_us_per_launch"] == pytest.approx(2000.0, rel=1e-3)
    # 40 B per cell per launch over 2 ms
    assert per["adj"]["achieved_GBs"] == pytest.approx(40 * cells / 2e-3 / 1e9, rel=1e-3)
    assert per["adj"]["frac"] == pytest.approx(per["adj"]["achieved_GBs"] / bench.HBM_PEAK_GBS, abs=1e-4)
    assert sum(v["share_of_time"] for v in per.values()) == pytest.approx(1.0, abs=1e-3)
    # one step per launch: the per-unit bytes of one cell-update
    one = bench.kernel_table({"fwd": (0.0, 0), "fwd_virt": (10.0, 100), "recomp": (0.0, 0), "adj": (20.0, 100)},
                             1000, 1)
    assert one["adj"]["bytes_per_cell_per_launch"] == 28 and one["fwd_virt"]["bytes_per_cell_per_launch"] == 20


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stderr
