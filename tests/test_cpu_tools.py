"""CPU tests of the measurement tools: the rocprofv3 summary (counter
calibration and per-launch traffic) on synthetic counter files, and the
bench's one-GPU guard."""
import csv
import json
import os
import subprocess
import sys

import pytest

from mrfmap_amd import profiling

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KIB = 1024


def _counters(path, name, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for i, (kernel, value) in enumerate(rows):
            w.writerow([i, kernel, name, value])


def test_summary_calibrates_counters_and_sums_kernel_families(tmp_path):
    d = tmp_path / "prof"
    cal = d / "cal"
    # the 1 GiB copy: FETCH_SIZE counts half of the bytes read, WRITE_SIZE all written
    _counters(str(cal / "fetch"), "FETCH_SIZE", [("__amd_rocclr_copyBuffer", 512 * KIB)])
    _counters(str(cal / "write"), "WRITE_SIZE", [("__amd_rocclr_copyBuffer", 1024 * KIB)])
    front = "void mrf::(anonymous namespace)::k_front_sums<0, 16>(int const*)"
    window = "void mrf::(anonymous namespace)::k_ray_window<0, 0, 8>(int const*)"
    other = "void mrf::(anonymous namespace)::k_front_sums<0, 8>(int const*)"
    # two launches of each kernel of the schedule in use, one of another schedule
    _counters(str(d / "fetch"), "FETCH_SIZE", [(front, 100.0), (front, 300.0), (window, 50.0), (window, 50.0),
                                               (other, 1e6)])
    _counters(str(d / "write"), "WRITE_SIZE", [(front, 10.0), (front, 10.0), (window, 40.0), (window, 40.0),
                                               (other, 1e6)])
    text, traffic = profiling.summarize(str(d), str(cal))
    assert "FETCH_SIZE x 2.000" in text and "WRITE_SIZE x 1.000" in text
    assert traffic["fetch_factor"] == pytest.approx(2.0) and traffic["write_factor"] == pytest.approx(1.0)
    by = {k["kernel"]: k for k in traffic["kernels"]}
    assert set(by) == {"k_front_sums<0, 16>", "k_ray_window<0, 0, 8>"}  # the family's most-launched variant
    # bytes per launch = 2 x mean FETCH_SIZE KiB x 1024 + mean WRITE_SIZE KiB x 1024
    assert by["k_front_sums<0, 16>"]["bytes"] == pytest.approx((2 * 200.0 + 10.0) * KIB)
    assert by["k_ray_window<0, 0, 8>"]["bytes"] == pytest.approx((2 * 50.0 + 40.0) * KIB)
    assert traffic["bytes"] == pytest.approx((410.0 + 140.0) * KIB)


def test_short_names_drop_namespaces_and_arguments():
    n = "void mrf::(anonymous namespace)::k_seg_sum<0>(unsigned long const*, float*)"
    assert profiling.short_name(n) == "k_seg_sum<0>"


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_committed_traffic_summary_is_consistent():
    path = os.path.join(ROOT, "profiles", "ray_kernels_real_13kf_f32_traffic.json")
    with open(path) as f:
        t = json.load(f)
    total = sum(k["bytes"] for k in t["kernels"])
    assert t["bytes"] == pytest.approx(total)
    for k in t["kernels"]:
        assert k["bytes"] == pytest.approx(k["fetch_kib"] * 1024 * t["fetch_factor"] + k["write_kib"] * 1024 * t["write_factor"])
