"""rocprofv3 CSV summaries behind the bench line's `roofline.traffic` and the
committed profiles: copy calibration, per-stage selection, counter averages.

The CSVs are synthetic, in the column layout rocprofv3 -f csv writes."""
import csv
import os

import pytest

from fastimagingpipe_amd import rocprof_summary as R

COPY = "__amd_rocclr_copyBuffer"
SPLAT = "void fip::k_splat<64>(float const*, fip::SnapMeta const*)"
C2R = "void fip::k_stage2<3072, 2>(HIP_vector_type<float, 2u> const*)"


def _counters(path, rows):
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "run_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for i, (kernel, counter, value) in enumerate(rows):
            w.writerow([i, i, kernel, counter, value])


def test_traffic_is_scaled_by_the_copy_calibration(tmp_path):
    n = 256 << 20
    # FETCH_SIZE reads half of a copy's bytes, WRITE_SIZE all of them (KiB).
    _counters(str(tmp_path / "f"), [(COPY, "FETCH_SIZE", n / 2048), (COPY, "FETCH_SIZE", 100.0),
                                    (SPLAT, "FETCH_SIZE", 1000.0), (SPLAT, "FETCH_SIZE", 3000.0)])
    _counters(str(tmp_path / "w"), [(COPY, "WRITE_SIZE", n / 1024), (SPLAT, "WRITE_SIZE", 500.0),
                                    (SPLAT, "WRITE_SIZE", 500.0)])
    t = R.traffic(str(tmp_path / "f"), str(tmp_path / "w"), n)
    c = t["calibration"]
    # the largest copy dispatch is the calibration copy, not the small upload
    assert c["fetch_factor"] == pytest.approx(2.0) and c["write_factor"] == pytest.approx(1.0)
    k = t["kernels"]["splat"]
    assert k["dispatches"] == 2
    assert k["fetch_bytes"] == pytest.approx(2000.0 * 1024 * 2.0)
    assert k["write_bytes"] == pytest.approx(500.0 * 1024)
    assert k["bytes"] == pytest.approx(k["fetch_bytes"] + k["write_bytes"])
    assert set(t["kernels"]) == {"splat"}


def test_traffic_needs_the_copy_kernel(tmp_path):
    _counters(str(tmp_path / "f"), [(SPLAT, "FETCH_SIZE", 1.0)])
    _counters(str(tmp_path / "w"), [(SPLAT, "WRITE_SIZE", 1.0)])
    with pytest.raises(RuntimeError):
        R.traffic(str(tmp_path / "f"), str(tmp_path / "w"), 1 << 20)


def test_counter_table_averages_per_dispatch_over_passes(tmp_path):
    _counters(str(tmp_path / "p1"), [(C2R, "SQ_WAVE_CYCLES", 100.0), (C2R, "SQ_WAVE_CYCLES", 300.0),
                                     (SPLAT, "SQ_WAVE_CYCLES", 50.0), (COPY, "SQ_WAVE_CYCLES", 7.0)])
    _counters(str(tmp_path / "p2"), [(C2R, "SQ_ACTIVE_INST_VALU", 20.0), (C2R, "SQ_ACTIVE_INST_VALU", 60.0)])
    tab = R.counter_table([str(tmp_path / "p1"), str(tmp_path / "p2")])
    assert set(tab) == {"row_c2r", "splat"}   # the copy kernel is no stage
    assert tab["row_c2r"]["SQ_WAVE_CYCLES"] == pytest.approx(200.0)
    assert tab["row_c2r"]["SQ_ACTIVE_INST_VALU"] == pytest.approx(40.0)
    assert tab["splat"] == {"SQ_WAVE_CYCLES": pytest.approx(50.0)}


def test_kernel_stats_are_sorted_by_total_time(tmp_path):
    d = tmp_path / "s"
    d.mkdir()
    with open(d / "run_kernel_stats.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage"])
        w.writerow([SPLAT, 25, 3.2e6, 1.28e5, 30.0])
        w.writerow([C2R, 25, 3.1e6, 1.24e5, 29.0])
        w.writerow([COPY, 6, 9.0e6, 1.5e6, 41.0])
    rows = R.kernel_stats(str(d))
    assert [r["name"] for r in rows] == [COPY, SPLAT, C2R]
    assert R.stage_of(rows[1]["name"]) == "splat" and R.stage_of(rows[0]["name"]) is None
    assert rows[2]["avg_ns"] == pytest.approx(1.24e5) and rows[2]["calls"] == 25


def test_trace_statistics_over_the_last_dispatches(tmp_path):
    """--last N averages each kernel's last N dispatches (dispatch order), so
    a warm-up dispatch before them does not count."""
    path = tmp_path / "trace"
    os.makedirs(path)
    with open(path / "run_kernel_trace.csv", "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kind", "Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        t = 0
        durs = {SPLAT: [200000, 110000, 120000, 100000], C2R: [90000, 100000]}
        for i in range(4):
            for k, d in durs.items():
                if i < len(d):
                    w.writerow(["KERNEL_DISPATCH", 10 * i + (k == C2R), k, t, t + d[i]])
                    t += d[i] + 1000
    rows = {r["name"]: r for r in R.kernel_trace_last(str(path), 3)}
    assert rows[SPLAT]["calls"] == 3
    assert rows[SPLAT]["avg_ns"] == pytest.approx(110000.0)
    assert rows[SPLAT]["median_ns"] == pytest.approx(110000.0)
    assert rows[SPLAT]["min_ns"] == 100000.0 and rows[SPLAT]["max_ns"] == 120000.0
    assert rows[C2R]["calls"] == 2 and rows[C2R]["avg_ns"] == pytest.approx(95000.0)
