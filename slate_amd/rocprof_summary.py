"""Summarise rocprofv3 CSV output for the library's kernels.

    python -m slate_amd.rocprof_summary trace DIR [--top N]
        per-kernel launches, total and average duration from a
        ``rocprofv3 --kernel-trace --output-format csv`` run
    python -m slate_amd.rocprof_summary pmc DIR COUNTER [--match TEXT] [--calib TEXT --calib-bytes B]
        per-kernel dispatches and average counter value per dispatch from a
        ``rocprofv3 --pmc COUNTER --output-format csv`` run; with --calib, the
        kernel whose name contains TEXT is a plain copy of B bytes each way and
        the reading is rescaled so that kernel reads (writes) exactly B bytes

DIR is searched recursively for ``*kernel_trace.csv`` / ``*counter_collection.csv``.
FETCH_SIZE and WRITE_SIZE are reported by the hardware in KiB per dispatch.
"""
import argparse
import collections
import csv
import glob
import os
import sys


def _files(root, suffix):
    out = sorted(glob.glob(os.path.join(root, "**", "*" + suffix), recursive=True))
    if not out:
        sys.exit("no *%s under %s" % (suffix, root))
    return out


def _col(row, *names):
    low = {k.lower(): k for k in row}
    for n in names:
        if n.lower() in low:
            return row[low[n.lower()]]
    raise KeyError("none of %s in %s" % (names, sorted(row)))


def short(name, width=90):
    name = name.replace("void ", "")
    return name if len(name) <= width else name[:width - 3] + "..."


def trace(root, top):
    stats = collections.defaultdict(lambda: [0, 0.0])
    for f in _files(root, "kernel_trace.csv"):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                name = _col(row, "Kernel_Name")
                t0 = int(_col(row, "Start_Timestamp"))
                t1 = int(_col(row, "End_Timestamp"))
                s = stats[name]
                s[0] += 1
                s[1] += (t1 - t0) * 1e-6
    total = sum(v[1] for v in stats.values())
    print("%-90s %8s %12s %10s %6s" % ("kernel", "calls", "total_ms", "avg_us", "pct"))
    for name, (calls, ms) in sorted(stats.items(), key=lambda kv: -kv[1][1])[:top]:
        print("%-90s %8d %12.3f %10.2f %6.2f" % (short(name), calls, ms, 1e3 * ms / calls, 100 * ms / total))
    print("total kernel time %.3f ms over %d launches" % (total, sum(v[0] for v in stats.values())))


def pmc(root, counter, match, calib, calib_bytes):
    per = collections.defaultdict(dict)  # kernel -> dispatch -> summed value
    for f in _files(root, "counter_collection.csv"):
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                if _col(row, "Counter_Name") != counter:
                    continue
                name = _col(row, "Kernel_Name")
                d = _col(row, "Dispatch_Id", "Correlation_Id")
                per[name][d] = per[name].get(d, 0.0) + float(_col(row, "Counter_Value"))
    scale = 1.0
    if calib:
        cands = [k for k in per if calib in k]
        if not cands:
            sys.exit("no kernel matching %r for the calibration" % calib)
        vals = [v for k in cands for v in per[k].values()]
        kib = max(vals)
        scale = (calib_bytes / 1024.0) / kib
        print("calibration: %s reads %.0f KiB for %.0f KiB moved -> scale %.4f"
              % (short(cands[0], 60), kib, calib_bytes / 1024.0, scale))
    print("%-90s %8s %16s %16s" % ("kernel", "disp", "avg_%s_KiB" % counter, "avg_MiB_scaled"))
    rows = []
    for name, d in per.items():
        if match and match not in name:
            continue
        avg = sum(d.values()) / len(d)
        rows.append((name, len(d), avg))
    for name, cnt, avg in sorted(rows, key=lambda r: -r[1] * r[2]):
        print("%-90s %8d %16.1f %16.3f" % (short(name), cnt, avg, avg * scale / 1024.0))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--top", type=int, default=25)
    p = sub.add_parser("pmc")
    p.add_argument("dir")
    p.add_argument("counter")
    p.add_argument("--match", default="")
    p.add_argument("--calib", default="")
    p.add_argument("--calib-bytes", type=float, default=0.0)
    a = ap.parse_args(argv)
    if a.cmd == "trace":
        trace(a.dir, a.top)
    else:
        pmc(a.dir, a.counter, a.match, a.calib, a.calib_bytes)


if __name__ == "__main__":
    main()
