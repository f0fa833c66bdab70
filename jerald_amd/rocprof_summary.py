"""Text summaries of rocprofv3 output for the LDL step benchmark.

Three commands, each reading the CSV files rocprofv3 wrote
(``--output-format csv``) under a directory, and printing a summary:

``stats DIR [--bench-json FILE]``
    the ``*_kernel_stats.csv`` of a ``--kernel-trace --stats`` run as a table
    (calls, average and total duration, share) with, when the bench's JSON
    line is given, the dominant kernel's HIP-event average next to rocprof's.

``traffic FETCH_DIR WRITE_DIR --n N --probe-bytes B``
    HBM bytes per launch of every kernel the benchmark times, from one
    ``--pmc FETCH_SIZE`` and one ``--pmc WRITE_SIZE`` run (KiB per dispatch),
    each scaled by what the same counter reads for the benchmark's plain
    device-to-device copy of B bytes (``__amd_rocclr_copyBuffer``).  The data
    lines are what ``bench.py`` reads:
    ``<timer name> <n> <dispatches> <fetch bytes> <write bytes>``.

``texture DIR [DIR ...]``
    the texture (vector memory) path of every timed kernel from ``--pmc``
    runs of the TA, TD and TCP counters below (``_sum`` over the CUs), each run
    with ``GRBM_GUI_ACTIVE``: busy and stall cycles as a share of CU cycles
    (256 CUs x GRBM_GUI_ACTIVE / 8, the counter summing the 8 XCDs), and the
    mean L2 read latency of the TCP's requests in cycles.

Kernels are named by the timer the library brackets them with (the names of
``ldl_timer_read``), so the summary and the benchmark's own per-kernel table
line up.
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys

# (substrings of the HIP kernel name, timer name); first match wins
_TIMERS = (
    (("k_zpass_c2r", "GammaBwdStore"), "zpass_c2r_gamma"),
    (("k_zpass_c2r", "LossStore"), "zpass_c2r_loss"),
    (("k_zpass_c2r", "RespBwdStore"), "zpass_c2r_resp"),
    (("k_zpass_c2r", "LossSq2Store"), "zpass_c2r_loss2"),
    (("k_zpass_c2r", "Tri3Store3"), "zpass_c2r3"),
    (("k_zpass_c2r",), "zpass_c2r"),
    (("k_zpass_r2c",), "zpass_r2c"),
    (("k_colpass",), "colpass"),
    (("k_bin_keys",), "paint_bin"),
    (("k_scan_sums",), "paint_scan"),
    (("k_scan_apply",), "paint_scan"),
    (("k_bin_scatter",), "paint_scatter"),
    (("k_paint_tiles<float, 1,",), "paint_deposit1"),
    (("k_paint_tiles<double, 1,",), "paint_deposit1"),
    (("k_paint_tiles<float, 3,",), "paint_deposit3"),
    (("k_paint_tiles<double, 3,",), "paint_deposit3"),
    (("k_mesh_absmax",), "paint_absmax"),
    (("k_paint_float",), "paint_deposit"),
    (("k_paint_fixed",), "paint_deposit"),
    (("k_fixed_to_float",), "paint_convert"),
    (("k_grad6",), "grad6"),
    (("k_div6",), "div6"),
    (("k_sm_x_fwd",), "sm_x_fwd"),
    (("k_sm_x_bwd",), "sm_x_bwd"),
    (("k_x_realfilter",), "x_realfilter"),
    (("k_gauss_line",), "gauss_line"),
    (("k_readout3_update",), "readout3_update"),
    (("k_sm_bwd_gather",), "sm_bwd_gather"),
    (("k_paint_vjp_pos",), "paint_vjp_pos"),
    (("k_loss_direct",), "loss"),
)
# timers over several kernels: dispatches per timed operation
_KERNELS_PER_OP = {"paint_scan": 2}
# timers whose three-component form launches the same kernel over three times
# the rows (the grid tells them apart)
_TRIPLE = {"zpass_r2c": "zpass_r2c3", "colpass": "colpass3"}
_PROBE = "__amd_rocclr_copyBuffer"


def timer_of(kernel):
    for keys, name in _TIMERS:
        if all(k in kernel for k in keys):
            return name
    return None


def _csvs(d, suffix):
    out = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not out:
        raise SystemExit(f"no *{suffix} under {d}")
    return out


def _short(name, width=110):
    return name if len(name) <= width else name[: width - 3] + "..."


def stats(d, bench_json=None, top=40):
    rows = []
    for f in _csvs(d, "kernel_stats.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                rows.append((int(r["Calls"]), float(r["AverageNs"]) / 1e3, float(r["TotalDurationNs"]) / 1e6,
                             float(r["Percentage"]), r["Name"]))
    rows.sort(key=lambda r: -r[2])
    lines = []
    if bench_json:
        with open(bench_json) as fh:
            b = json.loads([ln for ln in fh if ln.startswith("{")][-1])
        lines.append(f"# bench line of the same process: {b['ms_per_step']} ms/step over {b['steps']} timed steps "
                     f"({b['value']} {b['unit']})")
        rf = b.get("roofline") or {}
        if rf:
            k = rf["kernel"]
            match = [r for r in rows if timer_of(r[4]) == k and "<double" not in r[4]]
            if match:
                n = sum(r[0] for r in match)
                avg = sum(r[1] * r[0] for r in match) / n
                lines.append(f"# dominant kernel {k}: HIP-event average {1e3 * rf['avg_launch_ms']:.1f} us over the timed steps; "
                             f"rocprof average {avg:.1f} us over all {n} dispatches of the process")
    lines.append("")
    lines.append(f" {'calls':>5} {'avg_us':>10} {'total_ms':>10} {'pct':>6}  {'timer':<16} kernel")
    for c, avg, tot, pct, name in rows[:top]:
        lines.append(f" {c:5d} {avg:10.1f} {tot:10.3f} {pct:6.2f}  {timer_of(name) or '-':<16} {_short(name)}")
    if len(rows) > top:
        rest = rows[top:]
        lines.append(f" ... {len(rest)} more kernels, {sum(r[2] for r in rest):.3f} ms in total")
    return "\n".join(lines)


def _counter(d, counter):
    """{dispatch id: (kernel name, grid size, value)} of one counter run."""
    out = {}
    for f in _csvs(d, "counter_collection.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                key = (f, r["Dispatch_Id"])
                k, g, v = out.get(key, (r["Kernel_Name"], int(r.get("Grid_Size") or 0), 0.0))
                out[key] = (k, g, v + float(r["Counter_Value"]))
    return out


def _per_timer(disp):
    """{timer: [KiB per dispatch ...]} for the FP32 kernels, the three-component
    forms split off by grid size."""
    by = collections.defaultdict(list)
    grids = collections.defaultdict(set)
    for k, g, v in disp.values():
        t = timer_of(k)
        if t and "<double" not in k:
            grids[t].add(g)
            by[t].append((g, v))
    out = collections.defaultdict(list)
    for t, lst in by.items():
        g0 = min(grids[t])
        for g, v in lst:
            out[_TRIPLE[t] if t in _TRIPLE and g == 3 * g0 else t].append(v)
    return out


def _probe_kib(disp, probe_bytes, counter):
    vals = [v for k, g, v in disp.values() if _PROBE in k]
    want = probe_bytes / 2 / 1024.0  # one side of the copy in KiB
    # the probe copies twice; other device copies of the process are far smaller
    big = [v for v in vals if v > 0.25 * want]
    if not big:
        raise SystemExit(f"{counter}: no copy-probe dispatch ({_PROBE}) of about {want:.0f} KiB")
    return sum(big) / len(big), want


def traffic(fetch_dir, write_dir, n, probe_bytes, alg=None):
    fe, wr = _counter(fetch_dir, "FETCH_SIZE"), _counter(write_dir, "WRITE_SIZE")
    pf, want = _probe_kib(fe, probe_bytes, "FETCH_SIZE")
    pw, _ = _probe_kib(wr, probe_bytes, "WRITE_SIZE")
    ff, fw = want / pf, want / pw
    tf, tw = _per_timer(fe), _per_timer(wr)
    lines = [f"# calibration: bench.py's copy probe copies {probe_bytes // 2} bytes device to device ({_PROBE});",
             f"#   FETCH_SIZE reads {pf:.1f} KiB -> fetch factor {ff:.4f};  WRITE_SIZE reads {pw:.1f} KiB -> write factor {fw:.4f}",
             "# data lines (read by bench.py): <timer name> <n> <dispatches per op> <fetch bytes x factor> <write bytes x factor>",
             "# comment lines after each: raw KiB and the algorithmic bytes bench.py uses for the same kernel"]
    for t in sorted(set(tf) & set(tw)):
        kpo = _KERNELS_PER_OP.get(t, 1)
        nf, nw = len(tf[t]), len(tw[t])
        if nf != nw or nf % kpo:
            lines.append(f"# {t}: {nf} FETCH_SIZE and {nw} WRITE_SIZE dispatches, skipped")
            continue
        ops = nf // kpo
        rf = sum(tf[t]) / ops
        rw = sum(tw[t]) / ops
        fb, wb = rf * 1024 * ff, rw * 1024 * fw
        lines.append(f"{t} {n} {ops} {fb:.0f} {wb:.0f}")
        a = alg(t) if alg else None
        tail = f"algorithmic {a} bytes; traffic/algorithmic {(fb + wb) / a:.2f}" if a else "algorithmic - bytes"
        lines.append(f"#   raw fetch {rf:.1f} KiB, raw write {rw:.1f} KiB; {tail}")
    return "\n".join(lines)


# (column, counter): TA busy / address and data stalled by the texture cache
# (TC), TD busy / stalled by TC, TCP stalled on pending misses / tag conflicts
_TEX = [("TA busy", "TA_TA_BUSY_sum"), ("TA addrTC", "TA_ADDR_STALLED_BY_TC_CYCLES_sum"),
        ("TA dataTC", "TA_DATA_STALLED_BY_TC_CYCLES_sum"), ("TD busy", "TD_TD_BUSY_sum"),
        ("TD TCst", "TD_TC_STALL_sum"), ("TCP pend", "TCP_PENDING_STALL_CYCLES_sum"),
        ("TCP tagc", "TCP_READ_TAGCONFLICT_STALL_CYCLES_sum")]


def texture(dirs, ncu=256, nxcd=8):
    """Per timer: mean dispatch time, each _TEX counter as % of CU cycles (each
    run normalised by its own GRBM_GUI_ACTIVE) and the mean L2 read latency."""
    val = collections.defaultdict(lambda: collections.defaultdict(float))
    dur = collections.defaultdict(list)
    for i, d in enumerate(dirs):
        grbm = collections.defaultdict(float)
        sums = collections.defaultdict(lambda: collections.defaultdict(float))
        for f in _csvs(d, "counter_collection.csv"):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    t = timer_of(r["Kernel_Name"])
                    if not t or "<double" in r["Kernel_Name"]:
                        continue
                    c, v = r["Counter_Name"], float(r["Counter_Value"])
                    if c == "GRBM_GUI_ACTIVE":
                        grbm[t] += v
                        if i == 0:
                            dur[t].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
                    else:
                        sums[t][c] += v
        for t, cs in sums.items():
            cyc = ncu * grbm[t] / nxcd
            for c, v in cs.items():
                val[t][c] = 100.0 * v / cyc if cyc else 0.0
                val[t]["_raw_" + c] = v
    cols = [(h, c) for h, c in _TEX if any(c in v for v in val.values())]
    lines = ["# % of CU cycles; L2 lat = TCP_TCC_READ_REQ_LATENCY_sum / TCP_TCC_READ_REQ_sum (cycles per request)",
             f"{'timer':18s} {'us':>7s} " + " ".join(f"{h:>9s}" for h, _ in cols) + f" {'L2 lat':>7s}"]
    for t in sorted(val, key=lambda t: -(sum(dur[t]) / max(len(dur[t]), 1))):
        v = val[t]
        rq = v.get("_raw_TCP_TCC_READ_REQ_sum", 0.0)
        lat = v.get("_raw_TCP_TCC_READ_REQ_LATENCY_sum", 0.0) / rq if rq else float("nan")
        us = sum(dur[t]) / len(dur[t]) if dur[t] else float("nan")
        lines.append(f"{t:18s} {us:7.0f} " + " ".join(f"{v.get(c, float('nan')):9.1f}" for _, c in cols) + f" {lat:7.0f}")
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--bench-json")
    s.add_argument("--top", type=int, default=40)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--n", type=int, required=True)
    t.add_argument("--probe-bytes", type=int, required=True, help="bytes the copy probe moves (read + write)")
    x = sub.add_parser("texture")
    x.add_argument("dirs", nargs="+")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats(a.dir, a.bench_json, a.top))
    elif a.cmd == "texture":
        print(texture(a.dirs))
    else:
        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        sys.path.insert(0, root)
        import bench
        print(traffic(a.fetch_dir, a.write_dir, a.n, a.probe_bytes, alg=lambda k: bench.kernel_bytes(k, a.n, 4)))


if __name__ == "__main__":
    main()
