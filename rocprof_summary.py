"""Text summaries of rocprofv3 output for the measurement records.

Four commands:

  python rocprof_summary.py stats DIR --command "..." --out FILE
      per-kernel duration statistics of a ``rocprofv3 --kernel-trace`` run
      (SQLite ``*_results.db`` or ``*kernel_trace.csv`` under DIR).

  python rocprof_summary.py traffic --kernel REGEX \
      --fetch DIR --write DIR --copy-fetch DIR --copy-write DIR \
      --copy-kernel REGEX --copy-bytes B --command "..." --out FILE
      HBM bytes per launch of the kernels matching REGEX from two ``--pmc``
      passes (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by
      what the same counter reads for a plain copy kernel that moves B bytes in
      and B bytes out.  Writes ``bytes_per_launch: <int>`` among its lines.
      ``--fetch-above K`` keeps only the dispatches whose FETCH_SIZE exceeds
      K KiB (the same dispatch index in both passes).

  python rocprof_summary.py counters DIR --kernel REGEX --command "..." \
      [--trace DIR] --out FILE
      median per dispatch of every counter of a ``--pmc`` run, with the
      durations of a kernel-trace run of the same program beside them.

  python rocprof_summary.py device BENCH_JSON --command "..." --out FILE
      per-kernel device-clock times of the surrogate workload from the JSON
      line of ``bench.py --only mlp`` (no profiler involved).

Only the standard library is used; the module reads what rocprofv3 wrote and
never runs anything.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re
import sqlite3
import statistics


def _kernel_rows(path):
    """[(name, start_ns, end_ns)] from a kernel-trace directory."""
    dbs = sorted(glob.glob(os.path.join(path, "**", "*results.db"), recursive=True))
    rows = []
    if dbs:
        for db in dbs:
            con = sqlite3.connect(db)
            try:
                cols = [r[1] for r in con.execute("pragma table_info(kernels)")]
                name = "name" if "name" in cols else "kernel_name"
                rows += con.execute(f"select {name}, start, end from kernels").fetchall()
            finally:
                con.close()
        return rows
    for f in sorted(glob.glob(os.path.join(path, "**", "*kernel_trace.csv"), recursive=True)):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                rows.append((r["Kernel_Name"], int(r["Start_Timestamp"]), int(r["End_Timestamp"])))
    if not rows:
        raise SystemExit(f"no kernel trace found under {path}")
    return rows


def _pmc_rows(path, counter):
    """{kernel_name: [value per dispatch]} of one counter from a --pmc run."""
    out = {}
    files = sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True))
    if not files:
        raise SystemExit(f"no counter_collection.csv under {path} (run rocprofv3 with -f csv)")
    for f in files:
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                out.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    return out


def _q(ds, p):
    return ds[min(len(ds) - 1, int(p * len(ds)))]


def stats(args):
    rows = _kernel_rows(args.dir)
    by = {}
    for name, s, e in rows:
        by.setdefault(name, []).append((e - s) / 1000.0)
    lines = [f"# rocprofv3 kernel-trace summary: {args.title}", "",
             "command:", "", f"    {args.command}", "",
             "durations in microseconds (end - start of each dispatch)", "",
             f"{'kernel':62s} {'n':>7s} {'total_ms':>10s} {'mean':>9s} {'p10':>9s} {'p50':>9s} "
             f"{'p90':>9s} {'p99':>9s} {'max':>9s}"]
    total = sum(sum(v) for v in by.values())
    for name, ds in sorted(by.items(), key=lambda kv: -sum(kv[1]))[: args.top]:
        ds.sort()
        lines.append(f"{name[:62]:62s} {len(ds):7d} {sum(ds) / 1e3:10.3f} {statistics.mean(ds):9.2f} "
                     f"{_q(ds, .1):9.2f} {_q(ds, .5):9.2f} {_q(ds, .9):9.2f} {_q(ds, .99):9.2f} "
                     f"{ds[-1]:9.2f}")
    lines += ["", f"all kernels: {sum(len(v) for v in by.values())} dispatches, {total / 1e3:.3f} ms busy"]
    text = "\n".join(lines) + "\n"
    _emit(text, args.out)


def _per_dispatch(rows, regex):
    pat = re.compile(regex)
    vals = [v for name, vs in rows.items() if pat.search(name) for v in vs]
    if not vals:
        raise SystemExit(f"no dispatch of a kernel matching {regex!r}")
    vals.sort()
    return vals


def _durations(path, counter, regex, keep=None):
    """Dispatch durations (us) of the matching kernels from a --pmc csv that
    carries Start/End timestamps, in dispatch order (None without them)."""
    pat = re.compile(regex)
    out = []
    for f in sorted(glob.glob(os.path.join(path, "**", "*counter_collection.csv"), recursive=True)):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter or not pat.search(r["Kernel_Name"]):
                    continue
                if "Start_Timestamp" not in r or "End_Timestamp" not in r:
                    return None
                out.append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0)
    if keep is not None:
        out = [out[k] for k in keep]
    return out


def _dispatches(rows, regex):
    """Values of the kernels matching regex in dispatch order (unsorted)."""
    pat = re.compile(regex)
    return [v for name, vs in rows.items() if pat.search(name) for v in vs]


def traffic(args):
    fetch = _per_dispatch(_pmc_rows(args.fetch, "FETCH_SIZE"), args.kernel)
    write = _per_dispatch(_pmc_rows(args.write, "WRITE_SIZE"), args.kernel)
    keep = None
    if args.fetch_above is not None:
        # the dispatches whose FETCH_SIZE exceeds the threshold (the two
        # passes run the same program, so dispatch k is the same launch)
        f_all = _dispatches(_pmc_rows(args.fetch, "FETCH_SIZE"), args.kernel)
        w_all = _dispatches(_pmc_rows(args.write, "WRITE_SIZE"), args.kernel)
        if len(f_all) != len(w_all):
            raise SystemExit("the FETCH_SIZE and WRITE_SIZE passes saw different dispatch counts")
        keep = [k for k, v in enumerate(f_all) if v > args.fetch_above]
        if not keep:
            raise SystemExit("no dispatch above the FETCH_SIZE threshold")
        fetch = sorted(f_all[k] for k in keep)
        write = sorted(w_all[k] for k in keep)
    cf = _per_dispatch(_pmc_rows(args.copy_fetch, "FETCH_SIZE"), args.copy_kernel)
    cw = _per_dispatch(_pmc_rows(args.copy_write, "WRITE_SIZE"), args.copy_kernel)
    kib = 1024.0
    # counter scale: bytes the copy really moved / bytes the counter reports
    f_scale = args.copy_bytes / (statistics.median(cf) * kib)
    w_scale = args.copy_bytes / (statistics.median(cw) * kib)
    fm, wm = statistics.median(fetch), statistics.median(write)
    rd = fm * kib * f_scale
    wr = wm * kib * w_scale
    lines = [f"# rocprofv3 --pmc HBM traffic: {args.title}", "",
             "commands:", ""] + [f"    {c}" for c in args.command] + ["",
             f"kernel regex: {args.kernel}" + (f", dispatches with FETCH_SIZE above {args.fetch_above} KiB"
                                               if args.fetch_above is not None else ""),
             f"dispatches: FETCH_SIZE {len(fetch)}, WRITE_SIZE {len(write)}",
             f"FETCH_SIZE KiB per dispatch: median {fm:.1f} (p10 {_q(fetch, .1):.1f}, p90 {_q(fetch, .9):.1f})",
             f"WRITE_SIZE KiB per dispatch: median {wm:.1f} (p10 {_q(write, .1):.1f}, p90 {_q(write, .9):.1f})",
             "",
             f"copy calibration ({args.copy_kernel}, {args.copy_bytes} bytes read and written):",
             f"  FETCH_SIZE {statistics.median(cf):.1f} KiB -> scale {f_scale:.4f}",
             f"  WRITE_SIZE {statistics.median(cw):.1f} KiB -> scale {w_scale:.4f}",
             "",
             f"read_bytes_per_launch: {rd:.0f}",
             f"write_bytes_per_launch: {wr:.0f}",
             f"bytes_per_launch: {rd + wr:.0f}"]
    durs = _durations(args.fetch, "FETCH_SIZE", args.kernel, keep)
    if durs:
        dm = statistics.median(durs)
        lines += ["", "dispatch duration in the FETCH_SIZE pass (counter collection serialises dispatches):",
                  f"  median {dm:.2f} us over {len(durs)} dispatches -> {(rd + wr) / dm / 1e3:.1f} GB/s of HBM "
                  "traffic"]
    _emit("\n".join(lines) + "\n", args.out)


def counters(args):
    """Per-dispatch medians of every counter a --pmc run collected for the
    kernels matching the regex, with the dispatch durations of a kernel-trace
    run of the same program when --trace names one."""
    pat = re.compile(args.kernel)
    per = {}
    files = []
    for d in args.dir:
        found = sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True))
        if not found:
            raise SystemExit(f"no counter_collection.csv under {d}")
        files += found
    rows = []
    for f in files:
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                if pat.search(r["Kernel_Name"]):
                    rows.append((int(r["Dispatch_Id"]), r["Counter_Name"], float(r["Counter_Value"])))
    keep = None
    if args.select:
        # dispatches whose counter reaches the threshold; the passes run the
        # same program, so a dispatch id names the same launch in each
        name, lo = args.select.split(":")
        keep = {i for i, c, v in rows if c == name and v >= float(lo)}
    for i, c, v in rows:
        if keep is None or i in keep:
            per.setdefault(c, []).append(v)
    if not per:
        raise SystemExit(f"no dispatch of a kernel matching {args.kernel!r}")
    lines = [f"# rocprofv3 --pmc counters: {args.title}", "", "commands:" if len(args.command) > 1 else "command:", ""]
    lines += [f"    {c}" for c in args.command]
    lines += ["", f"kernel regex: {args.kernel}"]
    if args.select:
        lines += [f"dispatches with {args.select.replace(':', ' >= ')} only ({len(keep)})"]
    lines += ["",
             f"{'counter':34s} {'dispatches':>10s} {'median':>16s} {'p10':>16s} {'p90':>16s}"]
    for name in sorted(per):
        vs = sorted(per[name])
        lines.append(f"{name:34s} {len(vs):10d} {statistics.median(vs):16.1f} "
                     f"{_q(vs, .1):16.1f} {_q(vs, .9):16.1f}")
    if args.trace:
        ds = sorted((e - s) / 1000.0 for name, s, e in _kernel_rows(args.trace) if pat.search(name))
        if ds:
            lines += ["", f"kernel-trace durations ({args.trace_command or 'same program'}):",
                      f"  dispatches {len(ds)}, median {statistics.median(ds):.2f} us, "
                      f"mean {statistics.mean(ds):.2f} us"]
    _emit("\n".join(lines) + "\n", args.out)


def device_kernels(args):
    """Per-kernel device-clock times of the surrogate workload from the JSON
    line of ``bench.py --only mlp`` (md_kernel_time: every block's start and
    end on the GPU wall clock), with the network kernel's MFMA figures."""
    import json
    with open(args.json) as fh:
        d = json.loads(fh.read().strip().splitlines()[-1])
    if "mlp" not in d:
        d = d["other_workloads"]["mlp_surrogate"]
    lines = [f"# device-clock kernel times: {args.title}", "", "command:", "", f"    {args.command}", "",
             "times in microseconds per launch, first block start to last block end (span) and from the",
             "previous kernel's end (stream), from per-block s_memrealtime stamps (md_kernel_time)", ""]
    for leg in ("mlp", "lj_fp32"):
        r = d[leg]
        lines.append(f"{leg}: {r['ms_per_step'] * 1e3:.1f} us per step over the timed run, "
                     f"{r.get('rebuilds', '?')} rebuilds in {r.get('launches', '?')} launches")
        for k, v in r["kernels"].items():
            lines.append(f"  {k:14s} span {v['span_us']:8.1f}  stream {v['stream_us']:8.1f}  launches {v['launches']}")
        lines.append("")
    m = d["mlp"].get("k_mlp_net_mfma")
    if m:
        lines += ["k_mlp_net MFMA:",
                  f"  flops per launch {m['flops_per_launch']:.4g} (80 MFMA 16x16x32 BF16 per 16 atoms)",
                  f"  {m['achieved_tflops']:.1f} TFLOP/s in {m['span_us']:.1f} us = {m['frac']:.4f} of "
                  f"{m['peak_tflops_bf16_dense']:.0f} TFLOP/s dense BF16", ""]
    lines += [f"speed_vs_lj {d['speed_vs_lj']:.4f}, held-out force RMSE {d['heldout_force_rmse_rel']:.4f}, "
              f"largest periodic LJ check {d['mlp'].get('max_rmse_rel', float('nan')):.4f}"]
    _emit("\n".join(lines) + "\n", args.out)


def read_counters(path):
    """({counter: median per dispatch}, median kernel-trace duration in us or
    None) from a summary written by ``counters`` (None if absent)."""
    if not os.path.exists(path):
        return None
    vals, dur = {}, None
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) == 5 and parts[0].isupper() and parts[1].isdigit():
                vals[parts[0]] = float(parts[2])
            if line.strip().startswith("dispatches") and "median" in line:
                dur = float(line.split("median")[1].split("us")[0])
    return vals, dur


def read_rate(path):
    """(bytes per launch, median duration us) from a ``traffic`` summary with
    dispatch durations (None if absent)."""
    if not os.path.exists(path):
        return None
    b, dur = None, None
    with open(path) as fh:
        for line in fh:
            if line.startswith("bytes_per_launch:"):
                b = int(float(line.split(":", 1)[1]))
            if line.strip().startswith("median") and "dispatches" in line:
                dur = float(line.split()[1])
    return b, dur


def read_traffic(path):
    """bytes_per_launch from a summary written by ``traffic`` (None if absent)."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith("bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    return None


def _emit(text, out):
    if out:
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        with open(out, "w") as fh:
            fh.write(text)
    print(text, end="")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python rocprof_summary.py")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command", required=True)
    s.add_argument("--title", default="")
    s.add_argument("--top", type=int, default=25)
    s.add_argument("--out")
    t = sub.add_parser("traffic")
    t.add_argument("--kernel", required=True)
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-kernel", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--command", action="append", required=True)
    t.add_argument("--title", default="")
    t.add_argument("--fetch-above", type=float, default=None,
                   help="only dispatches whose FETCH_SIZE exceeds this many KiB")
    t.add_argument("--out")
    c = sub.add_parser("counters", help="per-dispatch medians of SQ counters and the dispatch durations")
    c.add_argument("dir", nargs="+", help="--pmc output directories (one per counter pass)")
    c.add_argument("--kernel", required=True)
    c.add_argument("--command", required=True, action="append",
                   help="the command of each pass (repeat once per pass)")
    c.add_argument("--title", default="")
    c.add_argument("--select", help="COUNTER:MIN: only dispatches whose COUNTER reaches MIN")
    c.add_argument("--trace", help="kernel-trace directory of the same program, for durations")
    c.add_argument("--trace-command")
    c.add_argument("--out")
    k = sub.add_parser("device", help="device-clock kernel times from a bench.py --only mlp JSON line")
    k.add_argument("json")
    k.add_argument("--command", required=True)
    k.add_argument("--title", default="")
    k.add_argument("--out")
    a = ap.parse_args(argv)
    {"stats": stats, "traffic": traffic, "counters": counters, "device": device_kernels}[a.cmd](a)


if __name__ == "__main__":
    main()
