"""CPU: bench.py's contract pieces that need no GPU -- more than one GPU is
refused with a message, and the calibrated HBM traffic it reports is the
figure recorded in the committed rocprofv3 summary for the run's launch
shape (none for another shape)."""

import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "one MI355X" in r.stderr
    assert r.stdout == ""


def test_traffic_figure_comes_from_the_profile_summary():
    import bench
    with open(os.path.join(ROOT, "profiles", "wave_traffic.md")) as f:
        text = f.read()
    rec = {m.group(1): (int(m.group(2)), m.group(3))
           for m in re.finditer(r"^(\w+) traffic_bytes_per_launch = (\d+) launch = (.+)$", text, re.M)}
    assert set(rec) == {"bwd_kernel", "fwd_kernel"}
    for k, (v, launch) in rec.items():
        assert bench.measured_traffic("wave", k, launch) == v
        # a run with another launch shape reports no figure
        assert bench.measured_traffic("wave", k, launch + " (other)") is None
    assert bench.measured_traffic("poisson", "bwd_kernel", rec["bwd_kernel"][1]) is None
    rec = {k: v for k, (v, _) in rec.items()}
    # every recorded figure is a table row's counters scaled by the copy calibration
    rows = [(m.group(1), float(m.group(2)), float(m.group(3)))
            for m in re.finditer(r"^\| (\w+)<[^|]*\|[^|]*\|[^|]*\| ([\d.]+) \| ([\d.]+) \|", text, re.M)]
    copy = re.search(r"^\| copy \(1 GiB in, 1 GiB out\) \|[^|]*\|[^|]*\| ([\d.]+) \| ([\d.]+) \|", text, re.M)
    f_fetch = 1024 * 1024 / float(copy.group(1))      # KiB fetched for 1 GiB read
    f_write = 1024 * 1024 / float(copy.group(2))
    assert abs(f_fetch - 1.99995) < 2e-4 and f_write == 1.0
    recs = re.findall(r"^(\w+) traffic_bytes_per_launch = (\d+) launch", text, re.M)
    assert len(recs) >= 2
    for k, v in recs:
        assert any(abs(round(fe * 1.99995 * 1024 + wr * f_write * 1024) - int(v)) <= 1 for kk, fe, wr in rows if kk == k), k


def test_launch_option_parses_every_shape():
    import bench
    assert bench.parse_launch("bins2,chunks512") == (("bins", 2), ("chunks", 512))
    assert bench.parse_launch("waves32, waves32") == (("waves", 32), ("waves", 32))
    for bad in ("bins2", "bins2,threads4", "waves,chunks5", "bins1,bins1,bins1"):
        try:
            bench.parse_launch(bad)
        except SystemExit:
            continue
        raise AssertionError(bad)


def test_default_steps_give_about_a_second_of_timed_work():
    """Per-configuration default step counts: the wave keeps the 100 steps
    the default run times; the others time about a second of their steps."""
    import bench
    assert bench.DEFAULT_STEPS["wave"] == (100, 10)
    step_ms = {"poisson": 0.42, "burgers": 0.17, "ho": 0.025, "inverse": 0.06}
    for cfg, ms in step_ms.items():
        k, w = bench.DEFAULT_STEPS[cfg]
        assert 0.3 <= k * ms / 1e3 <= 2.0, cfg
        assert 10 <= w <= k // 10, cfg
