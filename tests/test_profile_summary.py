"""The rocprofv3 summary tool on synthetic counter and trace files: the
copy-calibrated HBM bytes per finest-depth SpMV launch, and the kernel table
order (no GPU)."""
import csv
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _write_counters(path, spmv_kib, copy_kib):
    # dispatch order: a coarse SpMV, the finest depth's init, its working
    # SpMV launches, one launch that found the solve finished, the copy
    rows = [(1, "k_cg_spmv<4>", 77.0), (2, "k_cg_init", 10.0)]
    rows += [(3 + i, "void prg::k_cg_spmv<5>(...)", v) for i, v in enumerate(spmv_kib)]
    rows += [(10, "void prg::k_cg_spmv<5>(...)", 3.0), (11, "__amd_rocclr_copyBuffer", copy_kib)]
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Value"])
        for r in rows:
            w.writerow(r)


def test_traffic_scales_by_the_copy(tmp_path):
    fetch, write, out = tmp_path / "f.csv", tmp_path / "w.csv", tmp_path / "t.json"
    _write_counters(fetch, [1000.0, 1100.0], 262144.0)  # fetch reads half the copy: scale 2
    _write_counters(write, [400.0, 440.0], 524288.0)  # write reads all of it: scale 1
    r = subprocess.run([sys.executable, "-m", "poissonrecon_gpu_amd.profile_summary", "traffic", "--fetch",
                        str(fetch), "--write", str(write), "--workload", "synthetic", "--out", str(out)],
                       cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    d = json.loads(out.read_text())
    assert d["launches"] == 2  # the idle launch and the coarse depth's are left out
    assert d["copy_calibration"]["fetch_scale"] == 2.0 and d["copy_calibration"]["write_scale"] == 1.0
    assert d["hbm_fetch_bytes_per_launch"] == 1050 * 1024 * 2
    assert d["hbm_write_bytes_per_launch"] == 420 * 1024
    assert d["hbm_bytes_per_launch"] == 1050 * 1024 * 2 + 420 * 1024


def test_trace_table_by_total_time(tmp_path):
    stats = tmp_path / "s.csv"
    with open(stats, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"])
        w.writerow(["prg::(anonymous namespace)::k_small(int)", "3", "3000", "1000", "10.0"])
        w.writerow(["prg::(anonymous namespace)::k_big(int)", "2", "27000", "13500", "90.0"])
    r = subprocess.run([sys.executable, "-m", "poissonrecon_gpu_amd.profile_summary", "trace", "--stats", str(stats)],
                       cwd=ROOT, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[2] == "| k_big | 2 | 0.027 | 13.5 | 90.0 |"
    assert lines[3] == "| k_small | 3 | 0.003 | 1.0 | 10.0 |"
