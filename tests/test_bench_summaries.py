"""The bench's rocprof summarisers (host logic, synthetic counter files).

The roofline's traffic figures come from rocprofv3 counter CSVs: FETCH_SIZE
and WRITE_SIZE per dispatch, scaled by what the same counters read for a copy
kernel of known size.  These tests write small CSVs in rocprofv3's layout
and check the per-launch and per-cycle sums, the calibration and the cycle
windows (the cycles of one solve run from one k_bump to the next).
"""
import csv
import importlib
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def bench(tmp_path, monkeypatch):
    sys.path.insert(0, ROOT)
    b = importlib.import_module("bench")
    monkeypatch.setattr(b, "PROFILE_LARGE_TRAFFIC", str(tmp_path / "large.json"))
    monkeypatch.setattr(b, "PROFILE_TRAFFIC", str(tmp_path / "flux.json"))
    return b


COPY = "k_stream_copy(double2 const*, double2*, unsigned long)"
STAGE64 = "void mgcfd::k_flux_part_stage<double, 256, 6, 3, false, 5>(mgcfd::PartInfo const*)"
STAGE32 = "void mgcfd::k_flux_part_stage<float, 256, 6, 3, false, 6>(mgcfd::PartInfo const*)"


def one_cycle(stage):
    # name, workgroups, fetch KiB, write KiB, microseconds
    return [("mgcfd::k_calc_dt_seg(int4 const*)", 100, 10.0, 2.0, 5.0),
            (stage, 40, 100.0, 20.0, 30.0), (stage, 40, 100.0, 20.0, 30.0),
            ("mgcfd::k_down(int, int const*)", 100, 30.0, 10.0, 8.0),
            ("mgcfd::k_bump(int*, int const*)", 1, 0.0, 0.0, 2.0)]


def write_runs(tmp_path, cycles_fp64, cycles_mixed, fetch_scale=2.0):
    """A trace and two counter CSVs: the copy kernel, then two solves of
    'cycles' cycles each (the first cycle of each a warm-up), with setup
    kernels before each solve."""
    seq = [(COPY, 8192, (512 << 20) / 1024 / fetch_scale, (512 << 20) / 1024, 100.0)]
    for stage, n in ((STAGE64, cycles_fp64), (STAGE32, cycles_mixed)):
        seq.append(("mgcfd::k_fill5(int, double*)", 100, 1.0, 50.0, 4.0))     # setup, outside any cycle
        for _ in range(n):
            seq += one_cycle(stage)
    trace = tmp_path / "trace.csv"
    with open(trace, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Grid_Size_X", "Workgroup_Size_X", "Start_Timestamp",
                    "End_Timestamp"])
        t = 1000
        for d, (k, g, _, _, us) in enumerate(seq, 1):
            w.writerow([d, k, g * 256, 256, t, t + int(us * 1000)])
            t += int(us * 1000) + 500
    paths = []
    for name, col in (("FETCH_SIZE", 2), ("WRITE_SIZE", 3)):
        p = tmp_path / f"{name}.csv"
        with open(p, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["Dispatch_Id", "Kernel_Name", "Grid_Size", "Workgroup_Size", "Counter_Name", "Counter_Value"])
            for d, x in enumerate(seq, 1):
                w.writerow([d, x[0], x[1] * 256, 256, name, x[col]])
        paths.append(str(p))
    return str(trace), paths[0], paths[1]


def test_cycle_windows_skip_the_warm_up_and_the_setup(bench, tmp_path):
    trace, _, _ = write_runs(tmp_path, 4, 3)
    rows = []
    with open(trace, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], 0, 0.0))
    win, n = bench.cycle_windows(rows, r"k_flux_part_stage<double")
    assert n == 3 and len(win) == 3 * 5
    assert "k_calc_dt_seg" in win[0][1] and "k_bump" in win[-1][1]
    win, n = bench.cycle_windows(rows, r"k_flux_part_stage<float")
    assert n == 2 and len(win) == 2 * 5
    assert bench.cycle_windows(rows, r"no_such_kernel") == (None, 0)


def test_large_summary_bytes_and_time_per_cycle(bench, tmp_path, capsys):
    bench.summarize_large(*write_runs(tmp_path, 4, 3, fetch_scale=2.0))
    out = json.load(open(bench.PROFILE_LARGE_TRAFFIC))
    assert out["copy_kernel"] == {"fetch_scale": 2.0, "write_scale": 1.0}
    per_cycle_kib = sum(2.0 * f + w for _, _, f, w, _ in one_cycle(STAGE64))
    for v, n in (("part_atomic", 3), ("mixed_atomic", 2)):
        assert out[v]["cycles"] == n
        assert out[v]["bytes_per_cycle"] == round(per_cycle_kib * 1024)
        assert out[v]["kernel_us_per_cycle"] == pytest.approx(75.0)
        stage = [e for e in out[v]["kernels"] if "k_flux_part_stage" in e["kernel"]]
        assert len(stage) == 1 and stage[0]["dispatches_per_cycle"] == 2 and stage[0]["workgroups"] == 40
    assert "GB/s" in capsys.readouterr().out


def test_large_summary_refuses_runs_that_do_not_line_up(bench, tmp_path):
    trace, fetch, _ = write_runs(tmp_path, 4, 3)
    other = tmp_path / "other"
    other.mkdir()
    _, _, write = write_runs(other, 5, 3)        # one cycle more in the WRITE_SIZE run
    with pytest.raises(SystemExit):
        bench.summarize_large(trace, fetch, write)


def test_large_traffic_reads_the_committed_summary(bench, tmp_path):
    assert bench.large_traffic("part_atomic") is None
    with open(bench.PROFILE_LARGE_TRAFFIC, "w") as f:
        json.dump({"part_atomic": {"bytes_per_cycle": 123}}, f)
    assert bench.large_traffic("part_atomic") == 123
    assert bench.large_traffic("mixed_atomic") is None


def test_flux_summary_scales_by_the_copy_kernel(bench, tmp_path):
    rows = [(COPY, (512 << 20) / 1024 / 2.0, (512 << 20) / 1024),
            ("void mgcfd::k_flux_part<double, 256, 6, 3, false, 6>(x)", 1000.0, 500.0),
            ("void mgcfd::k_flux_part<double, 256, 6, 3, false, 6>(x)", 1200.0, 500.0)]
    paths = []
    for name, col in (("FETCH_SIZE", 1), ("WRITE_SIZE", 2)):
        p = tmp_path / f"{name}.csv"
        with open(p, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
            for d, r in enumerate(rows, 1):
                w.writerow([d, r[0], name, r[col]])
        paths.append(str(p))
    bench.summarize_pmc(*paths)
    out = json.load(open(bench.PROFILE_TRAFFIC))
    assert out["copy_kernel"]["fetch_scale"] == 2.0 and out["copy_kernel"]["write_scale"] == 1.0
    assert out["part_atomic"]["traffic_bytes"] == round((1100.0 * 2.0 + 500.0) * 1024)


def write_trace(path, launches):
    """A kernel-trace CSV: (name, workgroups, start us, end us) per dispatch."""
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Grid_Size_X", "Workgroup_Size_X", "Start_Timestamp",
                    "End_Timestamp"])
        for d, (k, g, a, b) in enumerate(launches, 1):
            w.writerow([d, k, g * 256, 256, int(a * 1000), int(b * 1000)])


def test_trace_bursts_split_on_idle_time(bench, tmp_path, capsys):
    flux = "void mgcfd::k_flux_part<double, 256, 6, 3, false, 6>(mgcfd::PartInfo const*)"
    # three back-to-back launches, 100 us idle, then two launches 1 us apart
    launches = [(flux, 1172, 0, 18), (flux, 1172, 18, 36), (flux, 1172, 36, 54),
                (flux, 1172, 154, 174), (flux, 1172, 175, 195)]
    write_trace(tmp_path / "t.csv", launches)
    bench.summarize_trace(str(tmp_path / "t.csv"), "bursts")
    rows = [l for l in capsys.readouterr().out.splitlines() if l.startswith("| 0") or l.startswith("| 1")]
    assert rows == ["| 0 | 3 | 18.00 | 18.00 | 0.00 |", "| 1 | 2 | 20.00 | 21.00 | 1.00 |"]


def test_trace_cycle_window_per_cycle_figures(bench, tmp_path, capsys):
    launches, t = [], 0.0
    for _ in range(2):                              # two cycles of 18 stage kernels and one k_down
        for _ in range(18):
            launches.append((STAGE64, 1172, t, t + 10.0))
            t += 10.0
        launches.append(("mgcfd::k_down(int, int const*)", 1172, t, t + 5.0))
        t += 5.0
    write_trace(tmp_path / "t.csv", launches[:-1])   # the window ends at the last stage kernel
    bench.summarize_trace(str(tmp_path / "t.csv"), "cycle")
    out = capsys.readouterr().out
    assert "36 fused stage kernels (2 cycles)" in out
    assert "| `void k_flux_part_stage<double, 256, 6, 3, false, 5>` | 1172 | 18.0 | 10.00 | 180.0 |" in out
    assert "| `k_down` | 1172 | 0.5 | 5.00 | 2.5 |" in out


def test_counter_means_of_the_matching_kernel(bench, tmp_path, capsys):
    flux = "void mgcfd::k_flux_part<double, 256, 6, 3, false, 6>(x)"
    with open(tmp_path / "c.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writerows([[1, flux, "SQ_WAVES", 4688], [2, flux, "SQ_WAVES", 4688], [1, flux, "SQ_INSTS_VALU", 100],
                     [2, flux, "SQ_INSTS_VALU", 300], [3, "k_down(int)", "SQ_WAVES", 7]])
    bench.summarize_trace(str(tmp_path / "c.csv"), "counters")
    out = capsys.readouterr().out
    assert "| SQ_WAVES | 4688 | 2 |" in out and "| SQ_INSTS_VALU | 200 | 2 |" in out
