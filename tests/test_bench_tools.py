"""The measurement helpers of bench.py that turn rocprofv3 output into the
roofline figures (CPU only: synthetic counter files with known answers)."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _write_csv(path, header, rows):
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(header)
        w.writerows(rows)


def test_pmc_fp64_counts_lanes_and_keeps_the_last_launches(tmp_path):
    # two dispatches per kernel; the scope had one launch in the profiled step,
    # so only the last dispatch of each kernel counts.  64 operations per ADD,
    # MUL or TRANS wave instruction, 128 per FMA.
    rows = []
    for d, (add, mul, fma, trans) in ((1, (1000, 1000, 1000, 1000)), (3, (10, 20, 30, 40))):
        for name, v in (("ADD", add), ("MUL", mul), ("FMA", fma), ("TRANS", trans)):
            rows.append([d, "adb::k_contact_c3r(adb::WorldArgs)", "SQ_INSTS_VALU_%s_F64" % name, v])
    for d, fma in ((2, 999), (4, 7)):
        rows.append([d, "adb::k_contact_psd9(adb::WorldArgs)", "SQ_INSTS_VALU_FMA_F64", fma])
        rows.append([d, "adb::k_contact_psd9(adb::WorldArgs)", "SQ_INSTS_VALU_ADD_F32", 12345])   # not FP64
    _write_csv(tmp_path / "c.csv", ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], rows)
    (tmp_path / "p.json").write_text(json.dumps({"contact_group": {"launches": 1, "ms": 0.1, "bytes": 0.0}}))
    out = bench.pmc_fp64(str(tmp_path / "c.csv"), str(tmp_path / "p.json"), ["contact_group"])
    e = out["scopes"]["contact_group"]
    assert e["launches"] == 1
    assert e["fp64_ops_per_launch"] == 64.0 * (10 + 20 + 40) + 128.0 * 30 + 128.0 * 7


def test_pmc_traffic_scales_by_the_copy_calibration(tmp_path):
    # the copy of M MB reads M MB: its FETCH_SIZE (KiB) fixes the scale; the
    # scope's kernels are averaged over their last N dispatches
    copy_mb = 64.0
    cal_kib = 2.0 * copy_mb * 1e6 / 1024.0     # a counter that reads twice the true bytes
    rows_f = [["__amd_rocclr_copyBuffer", cal_kib], ["adb::k_pair_reduce(x)", 100.0], ["adb::k_pair_reduce(x)", 300.0]]
    rows_w = [["__amd_rocclr_copyBuffer", copy_mb * 1e6 / 1024.0], ["adb::k_pair_reduce(x)", 50.0],
              ["adb::k_pair_reduce(x)", 70.0]]
    _write_csv(tmp_path / "f.csv", ["Kernel_Name", "Counter_Value"], rows_f)
    _write_csv(tmp_path / "w.csv", ["Kernel_Name", "Counter_Value"], rows_w)
    (tmp_path / "p.json").write_text(json.dumps({"asm_reduce": {"launches": 1, "ms": 0.1, "bytes": 1e6}}))
    out = bench.pmc_traffic(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), str(tmp_path / "p.json"), "asm_reduce",
                            copy_mb)
    assert abs(out["fetch_scale"] - 0.5) < 1e-12 and abs(out["write_scale"] - 1.0) < 1e-12
    assert abs(out["fetch_bytes"] - 1024.0 * 0.5 * 300.0) < 1e-6
    assert abs(out["write_bytes"] - 1024.0 * 70.0) < 1e-6
    assert out["algorithmic_bytes_per_launch"] == 1e6
    assert abs(out["traffic_bytes_per_launch"] - (out["fetch_bytes"] + out["write_bytes"])) < 1e-9


def test_kernel_stats_table(tmp_path):
    long_name = "void rocprim::detail::" + "x" * 120
    rows = [[long_name, 4, 8_000_000, 2_000_000, 75.0, 1_000_000, 3_000_000],
            ["adb::k_structure(adb::StructArgs)", 2, 2_000_000, 1_000_000, 25.0, 900_000, 1_100_000]]
    _write_csv(tmp_path / "s.csv", ["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage", "MinNs", "MaxNs"], rows)
    md = bench.kernel_stats_md(str(tmp_path / "s.csv"), "T", "cmd")
    assert "Total kernel time: 10.00 ms" in md
    assert "| `adb::k_structure(adb::StructArgs)` | 2 | 2.000 | 1000.0 | 900.0 | 1100.0 | 25.00 |" in md
    assert ("`" + long_name[:87] + "...`") in md


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
