"""bench.py's rocprofv3 summaries keep the dispatches between the first two
k_trace_mark launches (the headline run's timed steps) and drop everything
else (autotune candidates, training, later runs).  CPU only: synthetic CSVs in
rocprofv3's column layout."""
from __future__ import annotations

import csv
import json
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402

TRACE_COLS = ["Kind", "Agent_Id", "Queue_Id", "Stream_Id", "Thread_Id", "Dispatch_Id", "Kernel_Id", "Kernel_Name",
              "Correlation_Id", "Start_Timestamp", "End_Timestamp"]
PMC_COLS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size",
            "Kernel_Id", "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count",
            "Accum_VGPR_Count", "SGPR_Count", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]

# (kernel, duration us, fetch KiB, write KiB): autotune candidate, mark, two
# timed launches, mark, a later run's launch, mark, mark
DISPATCHES = [("void k_jacobi<1, 1>(SimConst)", 900.0, 9000.0, 900.0),
              ("k_trace_mark(unsigned int)", 1.0, 0.0, 0.0),
              ("void k_jacobi<1, 1>(SimConst)", 80.0, 100.0, 10.0),
              ("void k_jacobi<1, 1>(SimConst)", 90.0, 120.0, 12.0),
              ("void k_jacobi<1, 1>(SimConst)", 3.0, 1.0, 0.0),
              ("k_trace_mark(unsigned int)", 1.0, 0.0, 0.0),
              ("void k_jacobi<1, 1>(SimConst)", 500.0, 7000.0, 700.0),
              ("k_trace_mark(unsigned int)", 1.0, 0.0, 0.0),
              ("k_trace_mark(unsigned int)", 1.0, 0.0, 0.0)]


def _write(path, cols, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        w.writerows(rows)


def _trace_dir(tmp_path):
    d = tmp_path / "trace"
    t, rows = 1000, []
    for i, (name, us, _, _) in enumerate(DISPATCHES, 1):
        rows.append(dict(Kind="KERNEL_DISPATCH", Agent_Id="Agent 1", Queue_Id=1, Stream_Id=1, Thread_Id=1,
                         Dispatch_Id=i, Kernel_Id=1, Kernel_Name=name, Correlation_Id=i, Start_Timestamp=t,
                         End_Timestamp=t + int(us * 1000)))
        t += int(us * 1000) + 10
    _write(str(d / "c2_kernel_trace.csv"), TRACE_COLS, rows)
    _write(str(d / "c2_kernel_stats.csv"), ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage",
                                            "MinNs", "MaxNs", "StdDev"],
           [dict(Name="k_jacobi", Calls=5, TotalDurationNs=1, AverageNs=1, Percentage=100, MinNs=1, MaxNs=1,
                 StdDev=0)])
    return d


def _pmc_dir(tmp_path, name, counter, col):
    d = tmp_path / name
    rows = []
    for i, disp in enumerate(DISPATCHES, 1):
        rows.append({c: 0 for c in PMC_COLS} | dict(Dispatch_Id=i, Kernel_Name=disp[0], Counter_Name=counter,
                                                     Counter_Value=disp[col]))
    _write(str(d / "run_counter_collection.csv"), PMC_COLS, rows)
    return d


def _copy_dir(tmp_path, name, counter, kib):
    d = tmp_path / name
    row = {c: 0 for c in PMC_COLS} | dict(Dispatch_Id=1, Kernel_Name="__amd_rocclr_copyBuffer", Counter_Name=counter,
                                          Counter_Value=kib)
    _write(str(d / "run_counter_collection.csv"), PMC_COLS, [row])
    return d


def test_trace_summary_keeps_the_marked_window(tmp_path):
    out = tmp_path / "kernels.txt"
    assert bench.summarize_trace(str(_trace_dir(tmp_path)), str(out), "cmd") == 0
    text = out.read_text()
    row = next(line for line in text.splitlines() if line.startswith("k_jacobi"))
    calls, avg, kept, avg_kept = row.split()[1:5]
    assert (int(calls), int(kept)) == (3, 2)
    assert float(avg_kept) == pytest.approx(85.0, abs=0.01)
    assert float(avg) == pytest.approx((80 + 90 + 3) / 3, abs=0.01)
    assert "4 marks in the trace" in text


def test_pmc_summary_keeps_the_marked_window(tmp_path):
    f = _pmc_dir(tmp_path, "fetch", "FETCH_SIZE", 2)
    w = _pmc_dir(tmp_path, "write", "WRITE_SIZE", 3)
    cf = _copy_dir(tmp_path, "cf", "FETCH_SIZE", (1 << 30) / 1024 / 2)     # fetch counter reads half
    cw = _copy_dir(tmp_path, "cw", "WRITE_SIZE", (1 << 30) / 1024)
    out, js = tmp_path / "traffic.txt", tmp_path / "traffic.json"
    assert bench.summarize_pmc(str(f), str(w), str(cf), str(cw), str(out), str(js), "cmd", {}) == 0
    res = json.loads(js.read_text())
    assert res["fetch_scale"] == pytest.approx(2.0)
    # the 1 KiB launch is past convergence (under 2 % of the largest fetch): mean of 100 and 120 KiB
    fetch = (100 + 120) / 2 * 2.0 * 1024
    write = (10 + 12) / 2 * 1024
    assert res["bytes_per_launch"]["k_jacobi"] == pytest.approx(fetch + write, rel=1e-6)


def test_pmc_summary_needs_the_marks(tmp_path):
    d = tmp_path / "nomark"
    rows = [{c: 0 for c in PMC_COLS} | dict(Dispatch_Id=1, Kernel_Name="k_jacobi", Counter_Name="FETCH_SIZE",
                                            Counter_Value=5)]
    _write(str(d / "run_counter_collection.csv"), PMC_COLS, rows)
    cf = _copy_dir(tmp_path, "cf", "FETCH_SIZE", 1024)
    cw = _copy_dir(tmp_path, "cw", "WRITE_SIZE", 1024)
    with pytest.raises(RuntimeError, match="k_trace_mark"):
        bench.summarize_pmc(str(d), str(d), str(cf), str(cw), str(tmp_path / "o.txt"), None, "cmd", {})
