"""rocprofv3 CSV summaries: kernel table, per-dispatch counters and the copy-calibrated traffic figure."""

from __future__ import annotations

import csv
import os

from genesis_amd import profiling

COLS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _counters(d, name, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "pmc_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(COLS)
        for i, (kernel, value) in enumerate(rows):
            w.writerow([i, i, 1, kernel, name, value])


def test_traffic_scales_by_the_copy_calibration(tmp_path):
    gib = 1 << 30
    # the counter reads 2x the true bytes on this device: the copy of 1 GiB shows 2 GiB fetched and 2 GiB written
    _counters(tmp_path / "f", "FETCH_SIZE", [("rs::k_fused_wave<true, 0>", 2000.0), ("rs::k_fused_wave<true, 0>", 2200.0),
                                             ("vectorized_elementwise_kernel<4>", 2 * gib / 1024),
                                             ("vectorized_elementwise_kernel<4>", 3.0)])  # small: not a calibration
    _counters(tmp_path / "w", "WRITE_SIZE", [("rs::k_fused_wave<true, 0>", 500.0),
                                             ("vectorized_elementwise_kernel<4>", 2 * gib / 1024)])
    t = profiling.traffic(str(tmp_path / "f"), str(tmp_path / "w"), "k_fused_wave", gib)
    assert abs(t["FETCH_SIZE"]["factor"] - 0.5) < 1e-12
    assert t["FETCH_SIZE"]["kernel_dispatches"] == 2 and t["FETCH_SIZE"]["copy_dispatches"] == 1
    assert abs(t["bytes_per_launch"] - (2100.0 * 1024 * 0.5 + 500.0 * 1024 * 0.5)) < 1e-6


def test_kernel_stats_table(tmp_path):
    d = tmp_path / "t" / "host" / "123"
    os.makedirs(d)
    with open(d / "trace_kernel_stats.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs", "StdDev"])
        w.writerow(["small", 10, 1000, 100, 0.1, 90, 110, 1])
        w.writerow(["k_fused_wave", 200, 240_000_000, 1_200_000, 99.9, 1_100_000, 1_500_000, 10])
    s = profiling.kernel_stats(str(tmp_path / "t"))
    assert s[0]["name"] == "k_fused_wave" and s[0]["calls"] == 200 and abs(s[0]["avg_us"] - 1200.0) < 1e-9


def test_source_digest_is_stable():
    assert profiling.source_digest() == profiling.source_digest()


def test_valu_figures(tmp_path):
    d = tmp_path / "v"
    os.makedirs(d)
    vals = {"SQ_INSTS_VALU": 1000.0, "SQ_ACTIVE_INST_VALU": 400.0, "SQ_THREAD_CYCLES_VALU": 12800.0,
            "SQ_WAVE_CYCLES": 9000.0, "GRBM_GUI_ACTIVE": 100.0}
    with open(d / "pmc_counter_collection.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(COLS)
        for i, (c, v) in enumerate(vals.items()):
            w.writerow([0, 0, 1, "rs::k_fused_wave<true, 0>", c, v])
    v = profiling.valu(str(d), "k_fused_wave", simds=8, xcds=2)
    assert abs(v["valu_busy"] - 400.0 * 4 / (8 * 100.0 / 2)) < 1e-12
    assert abs(v["valu_lane_util"] - 12800.0 / (64 * 400.0)) < 1e-12
