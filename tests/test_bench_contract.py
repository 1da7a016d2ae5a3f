"""bench.py's contract on the CPU: one GPU only, and a committed rocprofv3 traffic summary for
every launch shape whose roofline line reports measured HBM bytes."""
import json
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


# the autotuned legs may land on any tiled shape: blocks 128/256/512, 1 or 2 records per thread
SHAPES = [f"{b}x{k}" for b in (128, 256, 512) for k in (1, 2)]
NAMES = (["traffic_color_tiled_cube128", "traffic_color_tiled_sphere96", "traffic_color_solve_cube128"]
         + [f"traffic_color_tiled_batched256_{s}" for s in SHAPES]
         + [f"traffic_color_tiled_corot_cube128_{s}" for s in SHAPES])


@pytest.mark.parametrize("name", NAMES)
def test_traffic_summary_present_and_consistent(name):
    """The summary's traffic_bytes_per_launch equals its own FETCH/WRITE arithmetic (KiB x the
    calibration factors) and is read back by bench.read_traffic."""
    import bench
    path = os.path.join(ROOT, "profiles", name + ".txt")
    text = open(path).read()
    got = bench.read_traffic(name)
    assert got is not None and got > 0
    m = re.search(r"traffic per launch = ([0-9.]+)\*([0-9.]+) \+ ([0-9.]+)\*([0-9.]+) KiB", text)
    assert m, "no 'traffic per launch' line"
    fetch, ff, write, wf = (float(v) for v in m.groups())
    cal = re.search(r"FETCH_SIZE ([0-9.]+) KiB,\s+WRITE_SIZE ([0-9.]+) KiB", text)
    assert cal, "no calibration line"
    exact_ff = 524288.0 / float(cal.group(1))  # a 512 MiB copy reads 524288 KiB
    assert abs(exact_ff - ff) < 1e-4 and abs(524288.0 / float(cal.group(2)) - wf) < 1e-4
    assert abs((fetch * exact_ff + write * wf) * 1024 - got) <= 1e-6 * got + 2
