"""The measurement tooling of bench.py on the host, with small rocprofv3-style
CSV files written here: per-dispatch counter means matched to the exact
kernel name, the copy-calibrated traffic file, the kernel-trace summary, and
the packed-id byte count of the split count's roofline."""
import argparse
import csv
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

ROWS = "wtc::k_split_rows<unsigned int, 8, 8, 1, 2>(wtc::GraphDesc const*, wtc::SplitDesc const*, unsigned long*, unsigned long long*)"
DECOY = "wtc::k_split_rows_extra<unsigned int>(wtc::GraphDesc const*)"
COPY = "wtc::k_copy_u32x4(uint4 const*, uint4*, unsigned long)"


def _counters(d, rows):
    """rows: (kernel, dispatch, counter, value); several rows of one dispatch
    (one per XCC) add up."""
    d.mkdir(parents=True, exist_ok=True)
    with open(d / "pmc_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for k, disp, c, v in rows:
            w.writerow({"Dispatch_Id": disp, "Kernel_Name": k, "Counter_Name": c, "Counter_Value": v})


def test_per_dispatch_sums_xcc_rows_and_matches_the_exact_kernel(tmp_path):
    _counters(tmp_path / "f", [(ROWS, 1, "FETCH_SIZE", 100), (ROWS, 1, "FETCH_SIZE", 50),
                               (ROWS, 2, "FETCH_SIZE", 130), (ROWS, 2, "FETCH_SIZE", 70),
                               (DECOY, 3, "FETCH_SIZE", 10 ** 9), (ROWS, 1, "WRITE_SIZE", 7)])
    mean, n = bench._per_dispatch(str(tmp_path / "f"), "k_split_rows", "FETCH_SIZE")
    assert n == 2 and mean == pytest.approx(175.0)  # (150 + 200) / 2; the decoy and WRITE_SIZE left out


def test_pmc_summary_scales_by_the_copy_calibration(tmp_path, monkeypatch):
    """A 1 MiB copy whose FETCH_SIZE reads 512 KiB and WRITE_SIZE 1024 KiB:
    fetch counts are doubled, writes kept."""
    _counters(tmp_path / "kf", [(ROWS, 1, "FETCH_SIZE", 1000), (ROWS, 2, "FETCH_SIZE", 3000)])
    _counters(tmp_path / "kw", [(ROWS, 1, "WRITE_SIZE", 10), (ROWS, 2, "WRITE_SIZE", 30)])
    _counters(tmp_path / "cf", [(COPY, 1, "FETCH_SIZE", 512)])
    _counters(tmp_path / "cw", [(COPY, 1, "WRITE_SIZE", 1024)])
    (tmp_path / "profiles").mkdir()
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    args = argparse.Namespace(pmc_summary=[str(tmp_path / x) for x in ("kf", "kw", "cf", "cw")],
                              pmc_kernel="k_split_rows", copy_calibration_mb=1, pmc_workload="rmat24",
                              spread=128, a=262144)
    bench.pmc_summary(args)
    out = json.load(open(tmp_path / "profiles" / "k_split_rows_rmat24_a262144_traffic.json"))
    assert out["fetch_scale"] == pytest.approx(2.0) and out["write_scale"] == pytest.approx(1.0)
    assert out["fetch_bytes"] == pytest.approx(2000 * 1024 * 2.0)
    assert out["write_bytes"] == pytest.approx(20 * 1024)
    assert out["bytes_per_launch"] == pytest.approx(out["fetch_bytes"] + out["write_bytes"])
    # the bench line finds the same file for any spread (the split count has none)
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    assert bench.load_traffic("rmat24", 12, 262144, "k_split_rows")["fetch_bytes"] == out["fetch_bytes"]


def test_trace_summary_per_kernel_times(tmp_path):
    d = tmp_path / "trace"
    d.mkdir()
    with open(d / "run_kernel_trace.csv", "w", newline="") as f:
        w = csv.DictWriter(f, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for s0, dur in ((0, 2_000_000), (10_000_000, 4_000_000)):
            w.writerow({"Kernel_Name": ROWS, "Start_Timestamp": s0, "End_Timestamp": s0 + dur})
        w.writerow({"Kernel_Name": COPY, "Start_Timestamp": 0, "End_Timestamp": 1_000_000})
    out = tmp_path / "summary.md"
    args = argparse.Namespace(trace_summary=str(d), summary_out=str(out), summary_title="t", summary_cmd="cmd")
    bench.trace_summary(args)
    text = out.read_text()
    row = [line for line in text.splitlines() if "k_split_rows" in line][0]
    cells = [c.strip() for c in row.strip("|").split("|")]
    assert cells[1] == "2" and float(cells[2]) == pytest.approx(3.0)  # calls, avg ms
    assert float(cells[3]) == pytest.approx(2.0) and float(cells[4]) == pytest.approx(4.0)
    assert cells[6] == "85.7%"  # 6 of 7 ms
    assert "cmd" in text


def test_split_id_bytes_follow_the_packing_layouts():
    """16/P bytes per tail id: P = 8 up to 2^16 top rows, 7 up to 2^18, 6 above."""
    assert bench.split_id_bytes(1 << 16) == 2.0
    assert bench.split_id_bytes((1 << 16) + 32) == pytest.approx(16 / 7)
    assert bench.split_id_bytes(1 << 18) == pytest.approx(16 / 7)
    assert bench.split_id_bytes((1 << 18) + 32) == pytest.approx(16 / 6)
    assert bench.split_id_bytes(1 << 20) == pytest.approx(16 / 6)
