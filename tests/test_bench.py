"""The measurement code of bench.py, on the CPU: the one-GPU guard, the
algorithmic byte model against its closed form, and the counter summaries on
synthetic rocprofv3 CSV files whose answers are known."""
import csv
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("mx,meqn,maux,expect", [(32, 3, 0, 54.375), (16, 3, 0, 61.5), (8, 3, 0, 78.0),
                                                  (16, 4, 0, 82.0), (16, 1, 2, 45.5)])
def test_algorithmic_bytes(mx, meqn, maux, expect):
    # read q with its halo, read aux with its halo, write q, FP64:
    # 8 [meqn (h + 1) + maux h] with h = (mx + 4)^2 / mx^2
    assert bench.algorithmic_bytes_per_cell(meqn, maux, mx) == pytest.approx(expect, rel=1e-12)


def _write_counters(path, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value",
                                          "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_pmc_summary_calibrates_against_the_copy(tmp_path):
    """FETCH_SIZE reporting half the copy's bytes gives a factor of 2; a step
    kernel's per-dispatch traffic is then fetch*2 + write."""
    kib = bench.COPY_BYTES / 1024
    copy = "__amd_rocclr_copyBuffer"
    step = "void hipclaw::step_tiles<hipclaw::RPShallow<float>, float, 16, 2, true, true, 4>(hipclaw::StepArgs)"
    fetch = [dict(Dispatch_Id=i, Kernel_Name=copy, Counter_Name="FETCH_SIZE", Counter_Value=kib / 2,
                  Start_Timestamp=0, End_Timestamp=1) for i in range(3)]
    fetch += [dict(Dispatch_Id=10 + i, Kernel_Name=step, Counter_Name="FETCH_SIZE", Counter_Value=1000.0,
                   Start_Timestamp=0, End_Timestamp=1) for i in range(2)]
    write = [dict(Dispatch_Id=i, Kernel_Name=copy, Counter_Name="WRITE_SIZE", Counter_Value=kib,
                  Start_Timestamp=0, End_Timestamp=1) for i in range(3)]
    write += [dict(Dispatch_Id=10 + i, Kernel_Name=step, Counter_Name="WRITE_SIZE", Counter_Value=500.0,
                   Start_Timestamp=0, End_Timestamp=1) for i in range(2)]
    _write_counters(str(tmp_path / "fetch" / "f_counter_collection.csv"), fetch)
    _write_counters(str(tmp_path / "write" / "w_counter_collection.csv"), write)
    out = bench.pmc_summary(str(tmp_path / "fetch"), str(tmp_path / "write"), cells=1024)
    assert "FETCH_SIZE factor 2.000, WRITE_SIZE factor 1.000" in out
    line = [l for l in out.splitlines() if l.startswith("kernel=void hipclaw::step_tiles")][0]
    traffic = (1000.0 * 2 + 500.0) * 1024
    assert "traffic_bytes=%.0f" % traffic in line
    assert "bytes_per_cell=%.3f" % (traffic / 1024) in line


def test_valu_summary_issue_share(tmp_path):
    """SQ_INSTS_VALU over a dispatch of known length: the issue share is
    VALU x 4 cycles / (SIMDs x cycles)."""
    step = "void hipclaw::step_tiles<hipclaw::RPShallow<float>, float, 16, 2, true, true, 4>(hipclaw::StepArgs)"
    dur_ns = 100000
    valu = 0.5 * bench.SIMDS * dur_ns * bench.CLOCK_GHZ / 4     # half the issue slots
    rows = [dict(Dispatch_Id=1, Kernel_Name=step, Counter_Name=n, Counter_Value=v, Start_Timestamp=0,
                 End_Timestamp=dur_ns)
            for n, v in (("SQ_INSTS_VALU", valu), ("SQ_WAVES", 3072), ("SQ_WAVE_CYCLES", 1000),
                         ("SQ_WAIT_ANY", 250))]
    _write_counters(str(tmp_path / "valu" / "v_counter_collection.csv"), rows)
    out = bench.valu_summary(str(tmp_path / "valu"), cells=2048)
    line = [l for l in out.splitlines() if l.startswith("kernel=")][0]
    assert "valu_issue_share=0.500" in line
    assert "wait_any_share=0.250" in line
    assert "valu_per_cell=%.2f" % (valu / 2048) in line


def test_counter_summary_shares(tmp_path):
    """Generic counters per cell, and wait counters as a share of wave cycles."""
    step = "void hipclaw::step_tiles<hipclaw::RPShallow<float>, float, 16, 2, true, true, 4>(hipclaw::StepArgs)"
    rows = [dict(Dispatch_Id=d, Kernel_Name=step, Counter_Name=n, Counter_Value=v, Start_Timestamp=0,
                 End_Timestamp=1)
            for d in (1, 2) for n, v in (("SQ_WAVE_CYCLES", 1000.0), ("SQ_WAIT_INST_LDS", 200.0),
                                         ("SQ_INSTS_VALU_FMA_F32", 4096.0))]
    _write_counters(str(tmp_path / "c" / "c_counter_collection.csv"), rows)
    out = bench.counter_summary(str(tmp_path / "c"), cells=1024)
    assert "SQ_WAIT_INST_LDS           per_dispatch=200 per_cell=0.195 share_of_wave_cycles=0.200" in out
    assert "SQ_INSTS_VALU_FMA_F32      per_dispatch=4096 per_cell=4.000" in out
