"""Summaries of rocprofv3 CSV output for the solver kernels.

    python -m eulersr_amd.profiling trace <dir>/<name>_kernel_trace.csv [--by-grid]
    python -m eulersr_amd.profiling pmc <fetch_counter_collection.csv> <write_counter_collection.csv>
    python -m eulersr_amd.profiling valu <counter_collection.csv>

``trace``: per kernel, the average duration over all dispatches and over the
last ``--tail`` dispatches (bench.py ends with back-to-back launches of each
stage kernel, the same launches it times with HIP events).
``pmc``: per kernel, FETCH_SIZE and WRITE_SIZE (KiB per dispatch), and the
same counters for a device copy of known size (bench.py --profile ends with
two 1 GiB copies), to give a calibration factor.
``valu``: per stage kernel, from one pass of SQ_WAVES, SQ_INSTS_VALU,
SQ_INSTS_VMEM_RD, SQ_WAVE_CYCLES and SQ_BUSY_CU_CYCLES, the means over the
last ``--tail`` dispatches of waves, VALU and vector-memory read
instructions per wave, and resident waves per busy CU.
"""
from __future__ import annotations

import argparse
import csv
import re
import statistics
from collections import defaultdict


def short(name: str) -> str:
    m = re.match(r"(?:void )?esr::(k_\w+)(<[^>]*>)?", name)
    return (m.group(1) + (m.group(2) or "")) if m else name[:40]


def trace_summary(path, tail=50, by_grid=False):
    """by_grid: one line per (kernel, grid size), which separates the
    shapes of one kernel (the GEMM's layers and passes)."""
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))

    def key(r):
        k = short(r["Kernel_Name"])
        if by_grid:
            k += "@" + "x".join(r.get(f"Grid_Size_{a}", "") for a in "XYZ")
        return k
    by = defaultdict(list)
    for r in rows:
        by[key(r)].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = []
    out.append(f"{'kernel':34s} {'calls':>7s} {'avg_us':>9s} {'tail_avg_us':>12s} {'grid':>10s}")
    grids = {}
    for r in rows:
        grids[key(r)] = r.get("Grid_Size_X", "")
    for k, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        t = v[-tail:] if len(v) >= tail else v
        out.append(f"{k:34s} {len(v):7d} {statistics.mean(v):9.2f} {statistics.mean(t):12.2f} {grids[k]:>10s}")
    return "\n".join(out)


def pmc_summary(fetch_csv, write_csv, copy_bytes=None, calib_kernel="__amd_rocclr_copyBuffer", tail=50,
                traffic_out=None, label=""):
    """Per-kernel FETCH_SIZE / WRITE_SIZE (KiB per dispatch): mean over all
    dispatches and over the last `tail` dispatches (bench.py ends with
    back-to-back launches of the stage kernels at the bench size). With
    copy_bytes, the largest dispatch of `calib_kernel` (a device copy of
    copy_bytes bytes each way) gives the factor from counter KiB to bytes;
    traffic_out then receives one line per kernel: name, tail fetch bytes,
    tail write bytes (calibrated)."""
    def load(path, counter):
        rows = [r for r in csv.DictReader(open(path)) if r.get("Counter_Name", counter) == counter]
        key = "Dispatch_Id" if rows and "Dispatch_Id" in rows[0] else None
        if key:
            rows.sort(key=lambda r: int(r[key]))
        vals = defaultdict(list)
        for r in rows:
            vals[short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return vals
    f = load(fetch_csv, "FETCH_SIZE")
    w = load(write_csv, "WRITE_SIZE")
    out = [f"{'kernel':34s} {'calls':>7s} {'FETCH_KiB':>11s} {'WRITE_KiB':>11s} {'tail_FETCH':>11s} "
           f"{'tail_WRITE':>11s}"]
    for k in sorted(set(f) | set(w)):
        fv = statistics.mean(f[k]) if f.get(k) else float("nan")
        wv = statistics.mean(w[k]) if w.get(k) else float("nan")
        ft = statistics.mean(f[k][-tail:]) if f.get(k) else float("nan")
        wt = statistics.mean(w[k][-tail:]) if w.get(k) else float("nan")
        out.append(f"{k:34s} {len(f.get(k, [])):7d} {fv:11.1f} {wv:11.1f} {ft:11.1f} {wt:11.1f}")
    ck = short(calib_kernel)
    if copy_bytes and f.get(ck) and w.get(ck):
        kib = copy_bytes / 1024
        fac_f = kib / max(f[ck])
        fac_w = kib / max(w[ck])
        out.append(f"calibration ({ck}, largest dispatch = {kib:.0f} KiB each way): counter KiB x "
                   f"{fac_f:.4f} (fetch), x {fac_w:.4f} (write) = bytes / 1024")
        if traffic_out:
            with open(traffic_out, "w") as fo:
                fo.write(f"# {label}\n")
                fo.write("# kernel  fetch_bytes  write_bytes  (mean of the last %d dispatches, calibrated)\n" % tail)
                for k in sorted(set(f) & set(w)):
                    fb = statistics.mean(f[k][-tail:]) * fac_f * 1024
                    wb = statistics.mean(w[k][-tail:]) * fac_w * 1024
                    fo.write(f"{k} {fb:.0f} {wb:.0f}\n")
    return "\n".join(out)


STAGE_KERNELS = ("k_recon", "k_cell", "k_face", "k_update")


def valu_summary(path, tail=50):
    """Executed VALU work of the stage kernels (see the module docstring)."""
    rows = list(csv.DictReader(open(path)))
    by = defaultdict(dict)
    name = {}
    for r in rows:
        d = int(r["Dispatch_Id"])
        name[d] = r["Kernel_Name"]
        by[d][r["Counter_Name"]] = float(r["Counter_Value"])
    per = defaultdict(list)
    for d in sorted(by):
        per[short(name[d])].append(by[d])
    out = [f"{'kernel':34s} {'waves':>8s} {'VALU/wave':>10s} {'VMEM_RD/wave':>13s} {'waves/busy CU':>14s}"]
    for k in sorted(per):
        if not k.split("<")[0] in STAGE_KERNELS:
            continue
        t = per[k][-tail:]
        waves = statistics.mean(c["SQ_WAVES"] for c in t)
        valu = statistics.mean(c["SQ_INSTS_VALU"] / c["SQ_WAVES"] for c in t)
        vmem = statistics.mean(c["SQ_INSTS_VMEM_RD"] / c["SQ_WAVES"] for c in t)
        occ = statistics.mean(c["SQ_WAVE_CYCLES"] / c["SQ_BUSY_CU_CYCLES"] for c in t)
        out.append(f"{k:34s} {waves:8.0f} {valu:10.1f} {vmem:13.1f} {occ:14.2f}")
    return "\n".join(out)


def read_traffic_label(path):
    """First header line of a traffic file (the run's launch configuration)."""
    with open(path) as f:
        first = f.readline()
    return first[1:].strip() if first.startswith("#") else ""


def read_traffic(path):
    """{kernel: (fetch_bytes, write_bytes)} from a pmc_summary traffic file."""
    out = {}
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        name, fb, wb = line.rsplit(None, 2)
        out[name] = (float(fb), float(wb))
    return out


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("csv")
    t.add_argument("--tail", type=int, default=50)
    t.add_argument("--by-grid", action="store_true")
    p = sub.add_parser("pmc")
    p.add_argument("fetch_csv")
    p.add_argument("write_csv")
    p.add_argument("--copy-bytes", type=float, default=None)
    p.add_argument("--calib-kernel", default="__amd_rocclr_copyBuffer")
    p.add_argument("--tail", type=int, default=50)
    p.add_argument("--traffic-out", default=None)
    p.add_argument("--label", default="")
    v = sub.add_parser("valu")
    v.add_argument("csv")
    v.add_argument("--tail", type=int, default=50)
    a = ap.parse_args()
    if a.cmd == "trace":
        print(trace_summary(a.csv, a.tail, a.by_grid))
    elif a.cmd == "valu":
        print(valu_summary(a.csv, a.tail))
    else:
        print(pmc_summary(a.fetch_csv, a.write_csv, a.copy_bytes, a.calib_kernel, a.tail, a.traffic_out, a.label))


if __name__ == "__main__":
    main()
