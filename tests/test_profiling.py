"""The rocprofv3 CSV summaries (eulersr_amd.profiling) on synthetic files of
known content: kernel-trace tail means, FETCH_SIZE / WRITE_SIZE calibration
against a copy of known size and the traffic file bench.py reads back, and
the VALU-per-wave summary of the stage kernels."""
import csv

import pytest

from eulersr_amd import profiling as PR

CELL = "void esr::k_cell<1, 100, 0>(esr::GroupDev const*, int)"
COPY = "__amd_rocclr_copyBuffer"


def _write(path, fields, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        w.writerows(rows)
    return path


def test_short_names():
    assert PR.short(CELL) == "k_cell<1, 100, 0>"
    assert PR.short("esr::k_recon(esr::GroupDev const*)") == "k_recon"
    assert PR.short(COPY) == COPY


def test_trace_tail_mean(tmp_path):
    rows = [dict(Kernel_Name=CELL, Start_Timestamp=1000 * i, End_Timestamp=1000 * i + (10_000 if i < 5 else 40_000),
                 Grid_Size_X=64) for i in range(10)]
    out = PR.trace_summary(_write(tmp_path / "t.csv", list(rows[0]), rows), tail=5)
    line = [x for x in out.splitlines() if x.startswith("k_cell")][0].split()
    # 10 calls, mean of all 25 us, mean of the last 5: 40 us
    assert line[-4:] == ["10", "25.00", "40.00", "64"]


def test_pmc_calibration_and_traffic_file(tmp_path):
    gib_kib = float(1 << 20)
    fetch = [dict(Dispatch_Id=0, Kernel_Name=COPY, Counter_Name="FETCH_SIZE", Counter_Value=gib_kib / 2),
             dict(Dispatch_Id=1, Kernel_Name=CELL, Counter_Name="FETCH_SIZE", Counter_Value=1000.0)]
    write = [dict(Dispatch_Id=0, Kernel_Name=COPY, Counter_Name="WRITE_SIZE", Counter_Value=gib_kib),
             dict(Dispatch_Id=1, Kernel_Name=CELL, Counter_Name="WRITE_SIZE", Counter_Value=500.0)]
    f = _write(tmp_path / "f.csv", list(fetch[0]), fetch)
    w = _write(tmp_path / "w.csv", list(write[0]), write)
    traffic = tmp_path / "traffic.txt"
    out = PR.pmc_summary(f, w, copy_bytes=1 << 30, traffic_out=traffic, label="launch cell,64,1")
    assert "x 2.0000 (fetch), x 1.0000 (write)" in out
    assert PR.read_traffic_label(traffic) == "launch cell,64,1"
    fb, wb = PR.read_traffic(traffic)["k_cell<1, 100, 0>"]
    assert fb == pytest.approx(2.0 * 1000 * 1024) and wb == pytest.approx(500 * 1024)


def test_valu_summary(tmp_path):
    rows = []
    for d in range(3):
        for c, v in (("SQ_WAVES", 100), ("SQ_INSTS_VALU", 60_000), ("SQ_INSTS_VMEM_RD", 2_300),
                     ("SQ_WAVE_CYCLES", 625), ("SQ_BUSY_CU_CYCLES", 100)):
            rows.append(dict(Dispatch_Id=d, Kernel_Name=CELL, Counter_Name=c, Counter_Value=v))
        rows.append(dict(Dispatch_Id=10 + d, Kernel_Name=COPY, Counter_Name="SQ_WAVES", Counter_Value=1))
    out = PR.valu_summary(_write(tmp_path / "v.csv", list(rows[0]), rows))
    lines = out.splitlines()
    assert len(lines) == 2                      # header + the stage kernel (the copy is not one)
    assert lines[1].split()[-4:] == ["100", "600.0", "23.0", "6.25"]
