"""Summaries of rocprofv3 output for the trace kernels (text, for the profiles/ folder).

  python -m graphic_lab_task_amd.profiling stats <dir> --command "<profiled command>"
      per-kernel calls / average / median / min duration from <dir>/*kernel_trace.csv
  python -m graphic_lab_task_amd.profiling traffic <fetch_dir> <write_dir> --workload W --command "..."
      HBM bytes per trace call from two --pmc runs (FETCH_SIZE, then WRITE_SIZE, KiB per
      dispatch); each counter is scaled by what it reads for the 256 MiB elementwise copy
      (torch.neg) that `bench.py --copy-calibration` runs in the same process: 256 MiB read /
      FETCH_SIZE of the copy, 256 MiB written / WRITE_SIZE of the copy.
"""
import argparse
import re
import csv
import glob
import os
import statistics
import sys
from collections import defaultdict

TRACE_KERNELS = ("k_trace_closest", "k_trace_tail")
COPY_BYTES = (64 << 20) * 4              # bytes read (and written) by the 256 MiB float copy
XCDS = 8                                  # MI355X: accelerator complex dies per GPU


def _rows(directory, suffix):
    files = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {directory}")
    with open(files[0], newline="") as fh:
        return list(csv.DictReader(fh))


def _col(row, *names):
    low = {k.lower(): k for k in row}
    for n in names:
        if n.lower() in low:
            return row[low[n.lower()]]
    raise KeyError(names)


def short_name(name):
    """Kernel name without return type, namespaces and parameter list (library kernels)."""
    name = name.replace("(anonymous namespace)::", "")
    m = re.search(r"\blrt::(k_\w+)", name)
    if m:
        return name[m.start(1):].split("(")[0]
    return name.split("(")[0][:80]


def kernel_stats(directory):
    by = defaultdict(list)
    for r in _rows(directory, "kernel_trace.csv"):
        name = _col(r, "Kernel_Name", "KernelName", "Name")
        dur = int(_col(r, "End_Timestamp", "End")) - int(_col(r, "Start_Timestamp", "Start"))
        by[short_name(name)].append(dur / 1e3)
    return by


def cmd_stats(a):
    by = kernel_stats(a.dir)
    out = [f"# rocprofv3 kernel trace summary", f"command: {a.command}", "",
           f"{'kernel':70s} {'calls':>6s} {'avg_us':>10s} {'median_us':>10s} {'min_us':>10s} {'total_us':>12s}"]
    for name, d in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        out.append(f"{name:70s} {len(d):6d} {statistics.mean(d):10.2f} {statistics.median(d):10.2f} "
                   f"{min(d):10.2f} {sum(d):12.1f}")
    print("\n".join(out))


def _counter_per_dispatch(directory, counter):
    """{short kernel name: [value per dispatch]} for one counter (summed over its instances)."""
    per = defaultdict(float)
    names = {}
    for r in _rows(directory, "counter_collection.csv"):
        if _col(r, "Counter_Name") != counter:
            continue
        did = _col(r, "Dispatch_Id")
        per[did] += float(_col(r, "Counter_Value"))
        names[did] = short_name(_col(r, "Kernel_Name"))
    by = defaultdict(list)
    for did, v in per.items():
        by[names[did]].append(v)
    return by


def cmd_traffic(a):
    fetch = _counter_per_dispatch(a.fetch_dir, "FETCH_SIZE")
    write = _counter_per_dispatch(a.write_dir, "WRITE_SIZE")
    copy_names = [k for k in fetch if "neg" in k.lower()]
    if not copy_names:
        raise SystemExit("no neg kernel in the FETCH_SIZE run (run bench.py with --copy-calibration)")
    ck = copy_names[0]
    kib_f, kib_w = statistics.median(fetch[ck]), statistics.median(write[ck])
    scale_f, scale_w = COPY_BYTES / (kib_f * 1024.0), COPY_BYTES / (kib_w * 1024.0)
    lines = [f"# HBM traffic per step from rocprofv3 counters", f"command: {a.command}",
             f"copy reference kernel: {ck}: reads and writes {COPY_BYTES} bytes; FETCH_SIZE {kib_f:.1f} KiB -> "
             f"scale {scale_f:.4f}, WRITE_SIZE {kib_w:.1f} KiB -> scale {scale_w:.4f}", ""]
    total = 0.0
    for k in (a.kernels.split(",") if a.kernels else TRACE_KERNELS):
        f = [v for name, vs in fetch.items() if name.startswith(k) and "true" not in name for v in vs]
        w = [v for name, vs in write.items() if name.startswith(k) and "true" not in name for v in vs]
        if not f:
            continue
        fb = statistics.median(f) * 1024.0 * scale_f
        wb = statistics.median(w) * 1024.0 * scale_w if w else 0.0
        total += fb + wb
        lines.append(f"{k:20s} dispatches {len(f):4d}  fetch {fb / 1e6:10.3f} MB  write {wb / 1e6:10.3f} MB (median)")
    lines += ["", f"workload: {a.workload}", f"hbm_bytes_per_step: {total:.0f}"]
    print("\n".join(lines))


def cmd_counters(a):
    """Median per-dispatch value of every counter found in the given runs, for one kernel, plus
    the ratios that say what bounds it."""
    vals = {}
    for d in a.dirs:
        names = sorted({_col(r, "Counter_Name") for r in _rows(d, "counter_collection.csv")})
        for c in names:
            per = _counter_per_dispatch(d, c)
            v = [x for k, xs in per.items() if k.startswith(a.kernel) and "true" not in k for x in xs]
            if v:
                vals[c] = statistics.median(v)
    lines = [f"# rocprofv3 counters, {a.kernel} (median per dispatch)", f"command: {a.command}", ""]
    lines += [f"{c:32s} {v:18.1f}" for c, v in sorted(vals.items())]
    derived = []
    if "TCC_HIT_sum" in vals and "TCC_MISS_sum" in vals:
        derived.append(("L2 hit rate", vals["TCC_HIT_sum"] / max(1.0, vals["TCC_HIT_sum"] + vals["TCC_MISS_sum"])))
    if "TA_BUSY_avr" in vals and "GRBM_GUI_ACTIVE" in vals:
        # GRBM_GUI_ACTIVE reads as the sum over the part's XCDs
        derived.append((f"TA busy / active cycles ({XCDS} XCDs)",
                        vals["TA_BUSY_avr"] * XCDS / max(1.0, vals["GRBM_GUI_ACTIVE"])))
    if "SQ_WAIT_INST_ANY" in vals and "SQ_WAVE_CYCLES" in vals:
        derived.append(("wave cycles waiting on an instruction",
                        vals["SQ_WAIT_INST_ANY"] / max(1.0, vals["SQ_WAVE_CYCLES"])))
    if derived:
        lines += ["", "derived:"] + [f"{k:44s} {v:8.3f}" for k, v in derived]
    print("\n".join(lines))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--workload", required=True)
    t.add_argument("--command", required=True)
    t.add_argument("--kernels", default="", help="comma-separated kernel name prefixes (default: the trace kernels)")
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--kernel", required=True, help="kernel name prefix")
    c.add_argument("--command", required=True)
    a = ap.parse_args(argv)
    {"stats": cmd_stats, "traffic": cmd_traffic, "counters": cmd_counters}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
