"""Summaries of rocprofv3 databases for the power path.

    python -m gtpower_amd.rocprof_summary trace DB [--out FILE] [--cmd "..."]
    python -m gtpower_amd.rocprof_summary pmc FETCH_DB WRITE_DB --copy FETCH_DB WRITE_DB --copy-mib N
                                          [--out FILE] [--json FILE] [--cmd "..."]

`trace` reads a `rocprofv3 --kernel-trace` database: per-kernel dispatch
statistics, the device VCD parse windows (token count to event scatter), and
the exact engine's evaluation windows -- from the start of a
full glitch annotation (glitch_annotate_kernel) to the end of the last power
kernel (power_cell_kernel / power_tile_kernel) before that run's
cycle_reduce_kernel -- which is the interval the engine times with HIP events.

`counters` reads `rocprofv3 --pmc` databases of one program (one per counter
pass) and reports, per kernel and per evaluation, the instruction mix and the
derived VALUBusy (100 SQ_ACTIVE_INST_VALU / CUs / GRBM_GUI_ACTIVE) and
VALUUtilization (100 SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)); with
--steps (a JSON of pin-event steps per evaluation by kernel name, written by
bench.py --steps-json) also VALU instructions per 64-lane step.

`tiles` reads one `rocprofv3 --pmc` database and reports the kernels of the
last tile-subset run (gtp_run_tiles, the surrogate's judge: from its
tile_index_kernel to its cycle_reduce_list_kernel) per dispatch, and their
wave VALU instructions per judged cycle beside a full run's (--full-valu
instructions per --full-cycles cycles).

`pmc` reads two `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` databases of
the same program, and two of plain device copies of a known size, and reports
HBM traffic per evaluation window: the counters (KiB per dispatch) summed over
the window's kernels, each scaled by known bytes / counter reading of the
copies.
"""
import argparse
import json
import sqlite3
import statistics
import sys

WINDOW_START = "glitch_annotate_kernel("
WINDOW_KERNELS = ("power_cell_kernel", "power_tile_kernel")
WINDOW_END = "cycle_reduce_kernel("


def _short(name):
    n = name[5:] if name.startswith("void ") else name
    return n.split("(")[0]


def dispatches(db):
    c = sqlite3.connect(db)
    rows = c.execute("select dispatch_id, name, start, end from kernels order by start").fetchall()
    return [dict(id=r[0], name=r[1], start=r[2], end=r[3]) for r in rows]


def windows(disp):
    """Evaluation windows: lists of dispatches, annotate first."""
    out, cur = [], None
    for d in disp:
        if WINDOW_START in d["name"]:
            cur = [d]
        elif cur is not None and any(k in d["name"] for k in WINDOW_KERNELS):
            cur.append(d)
        elif cur is not None and WINDOW_END in d["name"]:
            out.append(cur)
            cur = None
    return out


def kernel_table(disp):
    by = {}
    for d in disp:
        by.setdefault(_short(d["name"]), []).append((d["end"] - d["start"]) / 1e3)
    rows = sorted(by.items(), key=lambda kv: -sum(kv[1]))
    lines = [f"{'kernel':58s} {'calls':>6s} {'total_ms':>10s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s}"]
    for k, v in rows:
        lines.append(f"{k[:58]:58s} {len(v):6d} {sum(v) / 1e3:10.3f} {statistics.mean(v):10.1f} "
                     f"{min(v):10.1f} {max(v):10.1f}")
    return lines


VCD_START = "count_starts_kernel("
VCD_END = "scatter_kernel("


def vcd_windows(disp):
    """Device VCD parses: from the token count to the event scatter."""
    out, cur = [], None
    for d in disp:
        if VCD_START in d["name"]:
            cur = [d]
        elif cur is not None:
            cur.append(d)
            if VCD_END in d["name"]:
                out.append(cur)
                cur = None
    return out


def trace_summary(db):
    disp = dispatches(db)
    lines = ["per-kernel dispatch statistics (device time)"] + kernel_table(disp)
    vw = vcd_windows(disp)
    if vw:
        dur = [(w[-1]["end"] - w[0]["start"]) / 1e6 for w in vw]
        busy = [sum(d["end"] - d["start"] for d in w) / 1e6 for w in vw]
        lines += ["", f"device VCD parse windows (token count start -> event scatter end): {len(vw)}",
                  "  per window ms: " + " ".join(f"{x:.3f}" for x in dur),
                  "  kernel busy ms: " + " ".join(f"{x:.3f}" for x in busy),
                  f"  kernels per window: {len(vw[-1])}"]
        by = {}
        for d in vw[-1]:
            by.setdefault(_short(d["name"]), 0.0)
            by[_short(d["name"])] += (d["end"] - d["start"]) / 1e3
        lines.append("  last window by kernel (us): " +
                     ", ".join(f"{k.split('::')[-1]} {v:.0f}" for k, v in sorted(by.items(), key=lambda kv: -kv[1])))
    ws = windows(disp)
    if ws:
        dur = [(max(d["end"] for d in w) - w[0]["start"]) / 1e6 for w in ws]
        lines += ["", f"exact evaluation windows (annotate start -> last power kernel end): {len(ws)}",
                  "  per window ms: " + " ".join(f"{x:.3f}" for x in dur),
                  f"  mean {statistics.mean(dur):.3f} ms, min {min(dur):.3f} ms, max {max(dur):.3f} ms",
                  f"  kernels per window: {len(ws[0])}"]
    return lines, ws


def _counter_by_dispatch(db, counter):
    c = sqlite3.connect(db)
    rows = c.execute("select dispatch_id, kernel_name, sum(value) from counters_collection "
                     "where counter_name = ? group by dispatch_id order by dispatch_id", (counter,)).fetchall()
    return [(r[0], r[1], r[2]) for r in rows]


def _windows_from_counters(rows):
    out, cur = [], None
    for did, name, val in rows:
        if WINDOW_START in name:
            cur = [(did, name, val)]
        elif cur is not None and any(k in name for k in WINDOW_KERNELS):
            cur.append((did, name, val))
        elif cur is not None and WINDOW_END in name:
            out.append(cur)
            cur = None
    return out


def _copy_factor(db, counter, mib):
    rows = [r for r in _counter_by_dispatch(db, counter) if "copy" in r[1].lower() or "elementwise" in r[1].lower()]
    big = [v for _, _, v in rows if v > 0.5 * mib * 1024]
    if not big:
        raise SystemExit(f"{db}: no copy dispatch of about {mib} MiB found for {counter}")
    return mib * 1024.0 / statistics.median(big), statistics.median(big)


def pmc_summary(fetch_db, write_db, copy_fetch_db, copy_write_db, copy_mib):
    ff, f_read = _copy_factor(copy_fetch_db, "FETCH_SIZE", copy_mib)
    fw, w_read = _copy_factor(copy_write_db, "WRITE_SIZE", copy_mib)
    wf = _windows_from_counters(_counter_by_dispatch(fetch_db, "FETCH_SIZE"))
    ww = _windows_from_counters(_counter_by_dispatch(write_db, "WRITE_SIZE"))
    if not wf or not ww:
        raise SystemExit("no evaluation window found in the counter databases")
    fetch = [sum(v for _, _, v in w) * ff * 1024 for w in wf]
    write = [sum(v for _, _, v in w) * fw * 1024 for w in ww]
    per_kernel = {}
    for w in wf:
        for _, name, v in w:
            per_kernel.setdefault(_short(name), [0.0, 0.0])[0] += v * ff * 1024 / len(wf)
    for w in ww:
        for _, name, v in w:
            per_kernel.setdefault(_short(name), [0.0, 0.0])[1] += v * fw * 1024 / len(ww)
    res = {"fetch_bytes": statistics.mean(fetch), "write_bytes": statistics.mean(write),
           "traffic_bytes": statistics.mean(fetch) + statistics.mean(write),
           "windows": len(wf), "copy_mib": copy_mib,
           "fetch_scale": ff, "write_scale": fw,
           "copy_fetch_kib_read": f_read, "copy_write_kib_read": w_read}
    lines = [f"copy calibration: {copy_mib} MiB copy reads FETCH_SIZE {f_read:.0f} KiB, WRITE_SIZE {w_read:.0f} KiB "
             f"-> scale fetch x{ff:.4f}, write x{fw:.4f}",
             f"evaluation windows: {len(wf)} (fetch), {len(ww)} (write)",
             f"HBM traffic per evaluation window: fetch {res['fetch_bytes'] / 1e9:.3f} GB, "
             f"write {res['write_bytes'] / 1e9:.3f} GB, total {res['traffic_bytes'] / 1e9:.3f} GB",
             "", f"{'kernel':58s} {'fetch_GB':>10s} {'write_GB':>10s}  (per window)"]
    for k, (f, w) in sorted(per_kernel.items(), key=lambda kv: -kv[1][0]):
        lines.append(f"{k[:58]:58s} {f / 1e9:10.3f} {w / 1e9:10.3f}")
    return lines, res


def _counters_by_kernel(db):
    """{kernel: {counter: summed value}} and {kernel: dispatches}; counters
    with a GRBM_ prefix take the maximum over their instances per dispatch
    (elapsed cycles), the others the sum."""
    c = sqlite3.connect(db)
    rows = c.execute("select dispatch_id, kernel_name, counter_name, sum(value), max(value) "
                     "from counters_collection group by dispatch_id, counter_name").fetchall()
    by, disp = {}, {}
    for did, name, ctr, vsum, vmax in rows:
        k = _short(name)
        by.setdefault(k, {}).setdefault(ctr, 0.0)
        by[k][ctr] += vmax if ctr.startswith("GRBM_") else vsum
        disp.setdefault(k, set()).add(did)
    return by, {k: len(v) for k, v in disp.items()}


def counters_summary(dbs, steps=None, n_cu=256, n_xcd=8):
    agg, ndisp = {}, {}
    for db in dbs:
        by, nd = _counters_by_kernel(db)
        for k, d in by.items():
            for ctr, v in d.items():
                agg.setdefault(k, {})[ctr] = v   # one pass per counter set: later passes add counters
            ndisp[k] = max(ndisp.get(k, 0), nd.get(k, 0))
    n_eval = max(1, ndisp.get("gtp::glitch_annotate_kernel", 1))
    lines = [f"evaluations: {n_eval} (glitch_annotate_kernel dispatches); values per evaluation", "",
             f"{'kernel':34s} {'waves':>9s} {'VALU/wave':>9s} {'VALUBusy%':>9s} {'VALUUtil%':>9s} "
             f"{'LDS/VALU':>8s} {'VMEM/VALU':>9s} {'SALU/VALU':>9s} {'wait/wave':>9s} {'L2hit%':>7s} "
             f"{'steps':>10s} {'VALU/step':>9s}"]
    res = {}
    for k, d in sorted(agg.items(), key=lambda kv: -kv[1].get("SQ_INSTS_VALU", 0)):
        if "SQ_INSTS_VALU" not in d:
            continue
        g = lambda ctr: d.get(ctr, float("nan")) / n_eval
        waves, valu = g("SQ_WAVES"), g("SQ_INSTS_VALU")
        # GRBM_GUI_ACTIVE of a dispatch is summed over the XCDs: elapsed cycles = / n_xcd
        busy = 100.0 * g("SQ_ACTIVE_INST_VALU") / n_cu / (g("GRBM_GUI_ACTIVE") / n_xcd)
        util = 100.0 * g("SQ_THREAD_CYCLES_VALU") / (64.0 * g("SQ_ACTIVE_INST_VALU"))
        hit = 100.0 * g("TCC_HIT_sum") / (g("TCC_HIT_sum") + g("TCC_MISS_sum"))
        st = (steps or {}).get(k.replace("gtp::", ""))
        per_step = valu / (st / 64.0) if st else float("nan")
        res[k] = dict(waves=waves, valu=valu, valu_busy=busy, valu_util=util, lds=g("SQ_INSTS_LDS"),
                      vmem_rd=g("SQ_INSTS_VMEM_RD"), salu=g("SQ_INSTS_SALU"), wait_inst_any=g("SQ_WAIT_INST_ANY"),
                      l2_hit=hit, steps=st, valu_per_step=per_step)
        lines.append(f"{k.replace('gtp::', '')[:34]:34s} {waves:9.3g} {valu / waves:9.1f} {busy:9.1f} {util:9.1f} "
                     f"{g('SQ_INSTS_LDS') / valu:8.3f} {g('SQ_INSTS_VMEM_RD') / valu:9.3f} "
                     f"{g('SQ_INSTS_SALU') / valu:9.3f} {4 * g('SQ_WAIT_INST_ANY') / waves:9.3g} {hit:7.1f} "
                     f"{(st or float('nan')):10.4g} {per_step:9.1f}")
    lines += ["", "VALUBusy: 100 SQ_ACTIVE_INST_VALU (quad-cycles) / CUs / elapsed cycles (GRBM_GUI_ACTIVE / XCDs; "
              "kernels profiled one at a time); "
              "VALUUtil: active lanes per VALU cycle; wait/wave: SQ_WAIT_INST_ANY cycles per wave; "
              "VALU/step: wave VALU instructions per 64 pin-event steps"]
    return lines, res


def tiles_summary(db, judged_cycles, full_valu, full_cycles):
    c = sqlite3.connect(db)
    rows = c.execute("select dispatch_id, kernel_name, counter_name, sum(value) from counters_collection "
                     "group by dispatch_id, counter_name").fetchall()
    by, names = {}, {}
    for did, name, ctr, v in rows:
        by.setdefault(did, {})[ctr] = v
        names[did] = name
    ids = sorted(by)
    ends = [i for i in ids if "cycle_reduce_list_kernel" in names[i]]
    if not ends:
        return ["no tile-subset run in the database"], {}
    end = ends[-1]
    start = max(i for i in ids if i < end and "tile_index_kernel" in names[i])
    lines = [f"{'kernel':34s} {'waves':>7s} {'VALU/wave':>10s} {'SALU/wave':>10s} {'wait/wave':>10s} "
             f"{'cycles/wave':>12s}"]
    merge_valu = 0.0
    for i in ids:
        if i < start or i > end:
            continue
        d = by[i]
        w = max(d.get("SQ_WAVES", 0.0), 1.0)
        if any(k in names[i] for k in WINDOW_KERNELS):
            merge_valu += d.get("SQ_INSTS_VALU", 0.0)
        lines.append(f"{_short(names[i]).replace('gtp::', '')[:34]:34s} {w:7.0f} "
                     f"{d.get('SQ_INSTS_VALU', 0) / w:10.0f} {d.get('SQ_INSTS_SALU', 0) / w:10.0f} "
                     f"{d.get('SQ_WAIT_INST_ANY', 0) / w:10.0f} {d.get('SQ_WAVE_CYCLES', 0) / w:12.0f}")
    per_k = merge_valu / judged_cycles * 1000.0
    full_k = full_valu / full_cycles * 1000.0
    lines += ["", f"power kernels: {merge_valu:.3g} wave VALU instructions for {judged_cycles} judged cycles = "
                  f"{per_k:.3g} per 1000 cycles; a full run: {full_k:.3g} per 1000 cycles; ratio {per_k / full_k:.2f}"]
    return lines, {"merge_valu": merge_valu, "per_1000_cycles": per_k, "full_per_1000_cycles": full_k}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    t = sub.add_parser("trace")
    t.add_argument("db")
    p = sub.add_parser("pmc")
    p.add_argument("fetch_db")
    p.add_argument("write_db")
    p.add_argument("--copy", nargs=2, required=True, metavar=("FETCH_DB", "WRITE_DB"))
    p.add_argument("--copy-mib", type=int, required=True)
    p.add_argument("--json")
    k = sub.add_parser("counters")
    k.add_argument("dbs", nargs="+")
    k.add_argument("--steps", help="JSON: pin-event steps per evaluation by kernel name")
    k.add_argument("--json")
    j = sub.add_parser("tiles")
    j.add_argument("db")
    j.add_argument("--judged-cycles", type=int, required=True)
    j.add_argument("--full-valu", type=float, required=True)
    j.add_argument("--full-cycles", type=int, required=True)
    for q in (t, p, k, j):
        q.add_argument("--out")
        q.add_argument("--cmd", default="")
    a = ap.parse_args(argv)
    if a.mode == "trace":
        lines, _ = trace_summary(a.db)
    elif a.mode == "tiles":
        lines, _ = tiles_summary(a.db, a.judged_cycles, a.full_valu, a.full_cycles)
    elif a.mode == "counters":
        steps = json.load(open(a.steps)) if a.steps else None
        lines, res = counters_summary(a.dbs, steps)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(res, f, indent=1)
    else:
        lines, res = pmc_summary(a.fetch_db, a.write_db, a.copy[0], a.copy[1], a.copy_mib)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(res, f, indent=1)
    if a.cmd:
        lines = [f"command: {a.cmd}", ""] + lines
    text = "\n".join(lines) + "\n"
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    sys.stdout.write(text)


if __name__ == "__main__":
    main()
