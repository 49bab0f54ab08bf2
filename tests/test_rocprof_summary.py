"""The rocprofv3 summariser (host logic only) on small synthetic CSV files:
launches told apart by grid, counter scaling by the probe kernels, the MFMA
busy share."""
import csv
import json
import os
import subprocess
import sys

from conftest import ROOT

from hacapk_hmvm_amd import rocprof_summary as RS

A, B = "void hmvm::tile_kernel<double, false, 8, true, true, false>(args)", "void hmvm::probe_read_kernel(args)"


def write(path, fields, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        w.writerows(rows)


def test_trace_and_traffic(tmp_path):
    kt, fe, wr = (tmp_path / d for d in ("kt", "fe", "wr"))
    for d in (kt, fe, wr):
        d.mkdir()
    trace = [{"Kernel_Name": A, "Grid_Size_X": g, "Start_Timestamp": 0, "End_Timestamp": us * 1000}
             for g, us in ((1024, 200), (1024, 210), (512, 100))]
    trace.append({"Kernel_Name": B, "Grid_Size_X": 99, "Start_Timestamp": 0, "End_Timestamp": 180000})
    write(kt / "x_kernel_trace.csv", ["Kernel_Name", "Grid_Size_X", "Start_Timestamp", "End_Timestamp"], trace)
    cf = ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"]
    write(fe / "f_counter_collection.csv", cf, [
        {"Kernel_Name": B, "Grid_Size": 99, "Counter_Name": "FETCH_SIZE", "Counter_Value": 524288},
        {"Kernel_Name": A, "Grid_Size": 1024, "Counter_Name": "FETCH_SIZE", "Counter_Value": 1000},
        {"Kernel_Name": A, "Grid_Size": 512, "Counter_Name": "FETCH_SIZE", "Counter_Value": 300}])
    write(wr / "w_counter_collection.csv", cf, [
        {"Kernel_Name": "void hmvm::probe_copy_kernel(args)", "Grid_Size": 99, "Counter_Name": "WRITE_SIZE",
         "Counter_Value": 1048576},
        {"Kernel_Name": A, "Grid_Size": 1024, "Counter_Name": "WRITE_SIZE", "Counter_Value": 10},
        {"Kernel_Name": A, "Grid_Size": 512, "Counter_Name": "WRITE_SIZE", "Counter_Value": 20}])
    out = tmp_path / "t.json"
    subprocess.check_call([sys.executable, os.path.join(ROOT, "hacapk_hmvm_amd", "rocprof_summary.py"), "--trace",
                           str(kt), "--fetch", str(fe), "--write", str(wr), "--grid-a", "1024", "--grid-b", "512",
                           "--json", str(out), "--md", str(tmp_path / "t.md")])
    rec = json.load(open(out))
    assert rec["launch_us"] == {"A": 205.0, "B": 100.0}
    assert rec["fetch_scale"] == 2.0 and rec["write_scale"] == 1.0
    assert rec["hbm_bytes_A"] == 1024 * (1000 * 2.0 + 10) and rec["hbm_bytes_B"] == 1024 * (300 * 2.0 + 20)


def test_mfma_table(tmp_path):
    k = "void hmvm::tile_multi_kernel<double, false, 2, false>(args)"
    rows = [{"Kernel_Name": k, "Grid_Size": 64, "Counter_Name": n, "Counter_Value": v}
            for n, v in (("SQ_VALU_MFMA_BUSY_CYCLES", 1024 * 500), ("GRBM_GUI_ACTIVE", 8 * 1000),
                         ("SQ_INSTS_VALU_MFMA_MOPS_F64", 2e6))]
    write(tmp_path / "m_counter_collection.csv", ["Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"], rows)
    lines = RS.mfma_table(str(tmp_path))
    assert lines[-1] == "| `tile_multi_kernel<double, false, 2, false>` | 64 | 50.0 | 1.02 |"
