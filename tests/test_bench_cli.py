"""bench.py's host logic on the CPU: the one-GPU guard and the measured-traffic summary (counter
runs scaled by the copy calibration, keyed by the kernel sources' hash)."""
import csv
import json
import os
import subprocess
import sys

import pytest

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def _counter_run(d, counter, values):
    """A rocprofv3 --pmc output directory: one row per (kernel, dispatch value in KiB)."""
    os.makedirs(os.path.join(d, "host"), exist_ok=True)
    with open(os.path.join(d, "host", "run_counter_collection.csv"), "w", newline="") as fh:
        w = csv.DictWriter(fh, ["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for kern, vals in values.items():
            for v in vals:
                w.writerow({"Kernel_Name": kern, "Counter_Name": counter, "Counter_Value": v})


def _fake_summary_dir(tmp_path):
    d = tmp_path / "tr"
    probe = 1 << 30
    # the copy kernel: FETCH_SIZE reads half the bytes it moves, WRITE_SIZE all of them
    _counter_run(str(d / "copy_fetch"), "FETCH_SIZE", {"k_copy(uint4 const*, uint4*, unsigned long)": [probe / 2048.0]})
    _counter_run(str(d / "copy_write"), "WRITE_SIZE", {"k_copy(uint4 const*, uint4*, unsigned long)": [probe / 1024.0]})
    (d / "copy_fetch.log").write_text("noise\n" + json.dumps({"probe_bytes": probe}) + "\n")
    nodes = 1000
    fused = "void rblbm::k3_fused<rblbm::FormF64, 34, 15, 1>(rblbm::KParams)"
    flow, heat = "void rblbm::k3_flow<0>(rblbm::KParams)", "void rblbm::k3_thermal<false, 0>(rblbm::KParams)"
    # one-pass: 600 KiB counted fetch per step (two dispatches averaged), 300 KiB written
    _counter_run(str(d / "fetch"), "FETCH_SIZE", {fused: [500.0, 700.0], "k_init(x)": [9.0]})
    _counter_run(str(d / "write"), "WRITE_SIZE", {fused: [400.0, 400.0]})
    (d / "fetch.log").write_text(json.dumps({
        "config": {"nodes": nodes},
        "forms": {"fp64": {"form": {"name": "one-pass"}, "bytes_per_update": 464.0}}}) + "\n")
    # two-pass tag: the step is the sum of its two kernels
    _counter_run(str(d / "fetch_two_pass"), "FETCH_SIZE", {flow: [200.0], heat: [100.0]})
    _counter_run(str(d / "write_two_pass"), "WRITE_SIZE", {flow: [150.0], heat: [50.0]})
    (d / "fetch_two_pass.log").write_text(json.dumps({
        "config": {"nodes": nodes},
        "forms": {"fp64": {"form": {"name": "two-pass"}, "bytes_per_update": 464.0}}}) + "\n")
    return d, nodes


def test_traffic_summary_scales_counters_by_the_copy(tmp_path, monkeypatch):
    d, nodes = _fake_summary_dir(tmp_path)
    out = tmp_path / "traffic.json"
    monkeypatch.setattr(bench, "TRAFFIC", str(out))
    bench.traffic_summary(str(d))
    t = json.loads(out.read_text())
    assert t["kernel_sha"] == bench.kernel_sha() and t["nodes"] == nodes
    assert t["copy_calibration"]["fetch_factor"] == pytest.approx(2.0)
    assert t["copy_calibration"]["write_factor"] == pytest.approx(1.0)
    one = t["forms"]["fp64/one-pass"]
    assert one["fetch_bytes_per_step"] == pytest.approx(600.0 * 1024 * 2)   # k_init not counted
    assert one["write_bytes_per_step"] == pytest.approx(400.0 * 1024)
    assert one["ratio"] == pytest.approx((1200.0 + 400.0) * 1024 / (464.0 * nodes))
    two = t["forms"]["fp64/two-pass"]
    assert two["fetch_bytes_per_step"] == pytest.approx(300.0 * 1024 * 2)
    assert two["write_bytes_per_step"] == pytest.approx(200.0 * 1024)
    assert "k3_flow<0>" in two["kernel"] and "k3_thermal<false, 0>" in two["kernel"]

    # the bench line takes the entry of the form it ran, only for the sources it was measured on
    b, src = bench.measured_traffic("fp64", "one-pass")
    assert b == pytest.approx(one["hbm_bytes_per_step"]) and "traffic" in src
    assert bench.measured_traffic("fp32", "one-pass")[0] is None
    t["kernel_sha"] = "0" * 16
    out.write_text(json.dumps(t))
    b, src = bench.measured_traffic("fp64", "one-pass")
    assert b is None and "0000000000000000" in src


def test_committed_traffic_summary_has_every_default_form():
    with open(bench.TRAFFIC) as fh:
        t = json.load(fh)
    for key in ("fp64/one-pass", "fp32/one-pass", "fp16/two-pass", "fp64/two-pass"):
        e = t["forms"][key]
        assert e["algorithmic_bytes_per_step"] == pytest.approx(
            {"fp64": 464, "fp32": 232, "fp16": 128}[key.split("/")[0]] * t["nodes"])
        assert e["ratio"] == pytest.approx(e["hbm_bytes_per_step"] / e["algorithmic_bytes_per_step"])
        assert 1.0 <= e["ratio"] < 2.0
