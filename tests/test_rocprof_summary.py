"""The rocprofv3 summarizer on synthetic CSV files: kernel names map to the
timer names bench.py reads, the copy probe calibrates the byte counters, and
bench.py reads back the data lines it writes."""
import csv

import bench
from jerald_amd import rocprof_summary as R

_COLS = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Queue_Id", "Process_Id", "Thread_Id", "Grid_Size", "Kernel_Id",
         "Kernel_Name", "Workgroup_Size", "LDS_Block_Size", "Scratch_Size", "VGPR_Count", "Accum_VGPR_Count", "SGPR_Count",
         "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]


def _write(path, counter, rows):
    path.mkdir(parents=True)
    with open(path / "bench_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, _COLS)
        w.writeheader()
        for i, (name, grid, val) in enumerate(rows):
            # two rows per dispatch (two counter instances) that the summary adds up
            for part in (0.25, 0.75):
                w.writerow({c: 0 for c in _COLS} | {"Dispatch_Id": i, "Grid_Size": grid, "Kernel_Name": name,
                                                   "Counter_Name": counter, "Counter_Value": val * part})


def test_timer_names_cover_the_benchmark_kernels():
    assert R.timer_of("void ldl::k_zpass_c2r<float, 256, ldl::HalfRowLoad<float>, ldl::GammaBwdStore<float> >(long)") == "zpass_c2r_gamma"
    assert R.timer_of("void ldl::k_zpass_c2r<float, 256, ldl::HalfRowLoad<float>, ldl::RealRowStore<float> >(long)") == "zpass_c2r"
    assert R.timer_of("void ldl::k_paint_tiles<float, 3, 8, 16, 32, 512, true, true>(float const*)") == "paint_deposit3"
    assert R.timer_of("void ldl::k_paint_tiles<float, 1, 8, 16, 32, 512, true, false>(float const*)") == "paint_deposit1"
    assert R.timer_of("void at::native::vectorized_elementwise_kernel<4>") is None
    named = {name for _, name in R._TIMERS} | set(R._TRIPLE.values())
    assert named <= set(bench.KERNELS)


def test_traffic_calibration_and_bench_read_back(tmp_path, monkeypatch):
    probe = 2 * (64 << 20)  # 64 MiB copied: 64 MiB read + 64 MiB written
    kib = (64 << 20) / 1024
    gather = "void ldl::k_sm_bwd_gather<float, ldl::GeomP2<float> >(float const*)"
    col = "void ldl::k_colpass<float, 512, 1, ldl::PlainColOp<float> >(ldl::ColGeom)"
    fetch = [(R._PROBE, 100, kib / 2), (R._PROBE, 100, kib / 2), (R._PROBE, 10, 3.0),
             (gather, 1000, 3000.0), (gather, 1000, 5000.0), (col, 64, 100.0), (col, 192, 300.0),
             ("void ldl::k_sm_bwd_gather<double, ldl::GeomP2<double> >(double const*)", 1000, 1e9)]
    write = [(R._PROBE, 100, kib), (R._PROBE, 100, kib), (R._PROBE, 10, 3.0),
             (gather, 1000, 1000.0), (gather, 1000, 1000.0), (col, 64, 200.0), (col, 192, 600.0)]
    _write(tmp_path / "f" / "host" / "123", "FETCH_SIZE", fetch)
    _write(tmp_path / "w", "WRITE_SIZE", write)
    txt = R.traffic(str(tmp_path / "f"), str(tmp_path / "w"), 512, probe)
    assert "fetch factor 2.0000" in txt and "write factor 1.0000" in txt
    data = {ln.split()[0]: ln.split()[1:] for ln in txt.splitlines() if ln and not ln.startswith("#")}
    # the FP32 gathers only, averaged; FETCH_SIZE doubled by the probe's factor
    assert data["sm_bwd_gather"] == ["512", "2", str(round(4000 * 1024 * 2)), str(1000 * 1024)]
    # the launch over three times the rows is the three-component form
    assert data["colpass"][1] == "1" and data["colpass3"][1] == "1"
    assert data["colpass3"][2] == str(300 * 1024 * 2)
    f = tmp_path / "traffic.txt"
    f.write_text(txt + "\n")
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(f))
    assert bench.measured_traffic("sm_bwd_gather", 512) == 4000 * 1024 * 2 + 1000 * 1024
    assert bench.measured_traffic("sm_bwd_gather", 256) is None


def test_committed_traffic_summary_is_read_by_bench():
    # the committed summary covers the benchmark's dominant kernels at 512^3
    for k in ("sm_bwd_gather", "readout3_update", "paint_deposit3", "grad6"):
        t = bench.measured_traffic(k, 512)
        assert t is not None and t > 0


def test_texture_shares_of_cu_cycles(tmp_path):
    gather = "void ldl::k_sm_bwd_gather<float, ldl::GeomP2<float> >(float const*)"
    # one dispatch: GRBM_GUI_ACTIVE sums 8 XCDs of 1000 cycles; 256 CUs -> 256000 CU cycles
    d = tmp_path / "g"
    d.mkdir()
    with open(d / "pmc_counter_collection.csv", "w", newline="") as f:
        w = csv.DictWriter(f, _COLS)
        w.writeheader()
        row = {c: 0 for c in _COLS} | {"Dispatch_Id": 1, "Kernel_Name": gather, "Start_Timestamp": 1000, "End_Timestamp": 3000}
        for c, v in (("GRBM_GUI_ACTIVE", 8000.0), ("TD_TD_BUSY_sum", 128000.0), ("TCP_PENDING_STALL_CYCLES_sum", 64000.0),
                     ("TCP_TCC_READ_REQ_sum", 10.0), ("TCP_TCC_READ_REQ_LATENCY_sum", 6000.0)):
            w.writerow(row | {"Counter_Name": c, "Counter_Value": v})
    lines = R.texture([str(d)]).splitlines()
    head = lines[1].split()
    vals = dict(zip(["timer", "us", "TD busy", "TCP pend", "L2 lat"], lines[2].split()))
    assert "TD" in head and vals["timer"] == "sm_bwd_gather" and vals["us"] == "2"
    assert vals["TD busy"] == "50.0" and vals["TCP pend"] == "25.0" and vals["L2 lat"] == "600"
