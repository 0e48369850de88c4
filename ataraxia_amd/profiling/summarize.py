"""Summarise rocprofv3 result databases into short text tables.

usage: python ataraxia_amd/profiling/summarize.py <prof_dir> [--copy-bytes N]

<prof_dir> is the OUT directory of profile_1080p.sh (or a script of the same
shape): kt/ (kernel trace + stats), fetch/ and write/ (FETCH_SIZE and
WRITE_SIZE passes), copy_fetch/ and copy_write/ (the same counters on
copy_probe), sq/ (SQ counters).  Prints kernel durations, per-dispatch
counter means, and HBM bytes per dispatch scaled by the copy calibration
(bytes = counter * N / counter_reading_of_an_N_byte_copy).
"""
import argparse
import glob
import os
import sqlite3
from collections import defaultdict


def _db(d):
    files = sorted(glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True))
    return files[0] if files else None


def kernel_stats(path):
    cur = sqlite3.connect(path).cursor()
    rows = cur.execute("select name, total_calls, total_duration, average from top_kernels").fetchall()
    return rows


def counters(path):
    cur = sqlite3.connect(path).cursor()
    rows = cur.execute("select kernel_name, counter_name, value, dispatch_id from counters_collection").fetchall()
    agg = defaultdict(list)
    for k, c, v, _ in rows:
        agg[(k, c)].append(float(v))
    return {key: (sum(v) / len(v), len(v)) for key, v in agg.items()}


def short(name, width=70):
    return name if len(name) <= width else name[: width - 3] + "..."


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("prof_dir")
    ap.add_argument("--copy-bytes", type=int, default=268435456)
    args = ap.parse_args()
    d = args.prof_dir
    out = []
    kt = _db(os.path.join(d, "kt"))
    if kt:
        out.append("## kernel trace (rocprofv3 --kernel-trace --stats)")
        out.append(f"{'kernel':72s} {'calls':>6s} {'total_ms':>10s} {'avg_ms':>9s}")
        for name, calls, total, avg in kernel_stats(kt):  # the view reports microseconds
            out.append(f"{short(name, 72):72s} {calls:6d} {total / 1e3:10.3f} {avg / 1e3:9.4f}")
    cal = {}
    for cname, sub in (("FETCH_SIZE", "copy_fetch"), ("WRITE_SIZE", "copy_write")):
        p = _db(os.path.join(d, sub))
        if not p:
            continue
        for (k, c), (mean, n) in counters(p).items():
            if "copy_f4" in k and c == cname:
                cal[cname] = args.copy_bytes / mean  # bytes per counter unit
    if cal:
        out.append("")
        out.append("## calibration (copy_probe, %d-byte copy)" % args.copy_bytes)
        for c, f in cal.items():
            out.append(f"{c}: {args.copy_bytes / f:.1f} counter units per copy -> {f:.2f} bytes/unit")
    for sub in ("fetch", "write", "sq"):
        p = _db(os.path.join(d, sub))
        if not p:
            continue
        out.append("")
        out.append(f"## counters ({sub}), mean per dispatch")
        for (k, c), (mean, n) in sorted(counters(p).items()):
            if "rocclr" in k:
                continue
            extra = ""
            if c in cal:
                extra = f"  = {mean * cal[c] / 1e6:.3f} MB (calibrated)"
            out.append(f"{short(k, 60):60s} {c:22s} n={n:3d} mean={mean:14.1f}{extra}")
    p = _db(os.path.join(d, "sq"))
    if p:
        c = counters(p)
        for k in sorted({k for k, _ in c}):
            tc = c.get((k, "SQ_THREAD_CYCLES_VALU"))
            ai = c.get((k, "SQ_ACTIVE_INST_VALU"))
            if tc and ai and ai[0] > 0:
                out.append(f"{short(k, 60):60s} VALU utilisation (active lanes / 64) = "
                           f"{100.0 * tc[0] / (ai[0] * 64):.1f}%")
    print("\n".join(out))


if __name__ == "__main__":
    main()
