"""Summarises rocprofv3 CSV output of `bench.py` runs (kernel trace and the
FETCH_SIZE / WRITE_SIZE counter passes) into the short text kept under
profiles/ and the traffic record bench.py reads.

    python -m hacapk_hmvm_amd.rocprof_summary --trace DIR [--fetch DIR --write DIR]
        --grid-a THREADS --grid-b THREADS [--md OUT.md] [--json OUT.json] [--cmd "..."]
    python -m hacapk_hmvm_amd.rocprof_summary --mfma DIR [--md OUT.md] [--cmd "..."]

The two tile launches of one HMVM run the same kernel (tile_kernel, or
tile_multi_kernel for several right-hand sides); they are told apart by
their grid size (threads = 256 * ceil(tiles / 4) with 256-thread workgroups).
Counter values are KiB per dispatch; they are scaled by what the same counter
reads for the bandwidth-probe kernels of bench.py, which read (probe_read) and
write (probe_copy) exactly 1 GiB per dispatch.
"""
import argparse
import csv
import glob
import json
import os
import re
import statistics

GIB_KIB = 1024.0 * 1024.0


def short(name):
    m = re.match(r"(?:void )?([\w:]+)(<[^(]*>)?\(", name)
    if not m:
        return name[:60]
    base = m.group(1).split("::")[-1]
    return base + (m.group(2) or "")


def read_csv(d, suffix):
    files = glob.glob(os.path.join(d, "*" + suffix))
    if not files:
        raise SystemExit("no *%s in %s" % (suffix, d))
    with open(files[0]) as f:
        return list(csv.DictReader(f))


def trace_table(rows):
    """Durations (us) per (kernel, grid); grids are kept apart for the tile
    kernels only (launch A and B), other kernels are pooled (grid 0)."""
    by = {}
    for r in rows:
        k = short(r["Kernel_Name"])
        key = (k, int(r["Grid_Size_X"]) if k.startswith(("tile_kernel", "tile_multi_kernel")) else 0)
        by.setdefault(key, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    return by


def counters(rows, counter):
    by = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        key = (short(r["Kernel_Name"]), int(r["Grid_Size"]))
        by.setdefault(key, []).append(float(r["Counter_Value"]))
    return by


def mfma_table(d, nxcd=8, nsimd=1024):
    """MFMA busy share and MFMA FLOPs of the tile_multi_kernel launches from a
    --pmc run of SQ_VALU_MFMA_BUSY_CYCLES, SQ_INSTS_VALU_MFMA_MOPS_{F64,F32,
    BF16} and GRBM_GUI_ACTIVE: busy % = MFMA busy cycles summed over the SIMDs
    / (GRBM_GUI_ACTIVE per XCD x SIMDs); GRBM_GUI_ACTIVE is read summed over
    the XCDs.  FLOPs = MOPS x 512."""
    by = {}
    for r in read_csv(d, "_counter_collection.csv"):
        k = short(r["Kernel_Name"])
        if not k.startswith("tile_multi_kernel"):
            continue
        by.setdefault((k, int(r["Grid_Size"])), {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    lines = ["| kernel | grid | MFMA busy % | MFMA GFLOP per launch |", "|---|---|---|---|"]
    for (k, g), c in sorted(by.items()):
        m = {n: statistics.mean(v) for n, v in c.items()}
        busy = 100.0 * m["SQ_VALU_MFMA_BUSY_CYCLES"] / (m["GRBM_GUI_ACTIVE"] / nxcd * nsimd)
        mops = sum(m.get("SQ_INSTS_VALU_MFMA_MOPS_" + t, 0.0) for t in ("F64", "F32", "BF16"))
        lines.append("| `%s` | %d | %.1f | %.2f |" % (k, g, busy, mops * 512 / 1e9))
    return lines


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mfma", help="directory of a --pmc run with the MFMA counters (prints their table only)")
    ap.add_argument("--trace")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--grid-a", type=int, default=0)
    ap.add_argument("--grid-b", type=int, default=0)
    ap.add_argument("--storage-kernel", default="tile_kernel<double")
    ap.add_argument("--md")
    ap.add_argument("--json")
    ap.add_argument("--cmd", default="")
    ap.add_argument("--meta", default="{}", help="JSON merged into the traffic record (N, len, ...)")
    a = ap.parse_args()
    if a.mfma:
        lines = ["# rocprofv3 MFMA counters: %s" % a.cmd, ""] + mfma_table(a.mfma)
        if a.md:
            with open(a.md, "w") as f:
                f.write("\n".join(lines) + "\n")
        else:
            print("\n".join(lines))
        return
    if not a.trace:
        ap.error("--trace is required")

    tr = trace_table(read_csv(a.trace, "_kernel_trace.csv"))
    lines = ["# rocprofv3 kernel trace: %s" % a.cmd, "", "| kernel | grid | calls | avg us | min us | max us |",
             "|---|---|---|---|---|---|"]
    for (k, g), d in sorted(tr.items(), key=lambda kv: -sum(kv[1]))[:14]:
        lines.append("| `%s` | %d | %d | %.1f | %.1f | %.1f |" % (k, g, len(d), statistics.mean(d), min(d), max(d)))

    launch = {}
    for tag, grid in (("A", a.grid_a), ("B", a.grid_b)):
        d = [v for (k, g), vs in tr.items() if k.startswith(a.storage_kernel) and g == grid for v in vs]
        if d:
            launch[tag] = statistics.mean(d)
    rec = dict(json.loads(a.meta))
    rec["launch_us"] = launch
    if a.fetch and a.write:
        fe = counters(read_csv(a.fetch, "_counter_collection.csv"), "FETCH_SIZE")
        wr = counters(read_csv(a.write, "_counter_collection.csv"), "WRITE_SIZE")
        f_probe = [v for (k, _), vs in fe.items() if k.startswith("probe_read_kernel") for v in vs]
        w_probe = [v for (k, _), vs in wr.items() if k.startswith("probe_copy_kernel") for v in vs]
        f_scale = GIB_KIB / statistics.mean(f_probe)
        w_scale = GIB_KIB / statistics.mean(w_probe)
        rec.update({"fetch_scale": f_scale, "write_scale": w_scale})
        lines += ["", "# HBM traffic (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE)", "",
                  "Calibration: probe_read_kernel reads 1 GiB per dispatch, FETCH_SIZE reads %.0f KiB "
                  "(scale %.3f); probe_copy_kernel writes 1 GiB, WRITE_SIZE reads %.0f KiB (scale %.3f)." %
                  (statistics.mean(f_probe), f_scale, statistics.mean(w_probe), w_scale), "",
                  "| launch | grid | FETCH KiB | WRITE KiB | HBM bytes (scaled) |", "|---|---|---|---|---|"]
        for tag, grid in (("A", a.grid_a), ("B", a.grid_b)):
            f = [v for (k, g), vs in fe.items() if k.startswith(a.storage_kernel) and g == grid for v in vs]
            w = [v for (k, g), vs in wr.items() if k.startswith(a.storage_kernel) and g == grid for v in vs]
            if not f or not w:
                continue
            hbm = 1024.0 * (statistics.mean(f) * f_scale + statistics.mean(w) * w_scale)
            rec["hbm_bytes_" + tag] = hbm
            lines.append("| %s | %d | %.0f | %.0f | %.4g |" % (tag, grid, statistics.mean(f), statistics.mean(w), hbm))
    if a.md:
        with open(a.md, "w") as f:
            f.write("\n".join(lines) + "\n")
    else:
        print("\n".join(lines))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(rec, f, indent=1)


if __name__ == "__main__":
    main()
