"""Summaries of rocprofv3 result databases (rocpd SQLite).

    python -m dmd_amd.profiling stats DB [--top N]
        kernel table (calls, total, average duration) from a
        `rocprofv3 --kernel-trace --stats` run, as markdown

    python -m dmd_amd.profiling window DB --steps N [--anchor KERNEL]
        per-kernel time per step over the last N steps of a kernel trace (the
        span from the N-th last dispatch of the anchor kernel, one per step,
        to the end of its last one), with each kernel's stream, grid,
        workgroup size, VGPRs and LDS: the steady state after warmup

    python -m dmd_amd.profiling counters DB [--kernel NAME ...] [--last N]
        mean value per dispatch of every counter of a `rocprofv3 --pmc` run,
        per kernel, as markdown

    python -m dmd_amd.profiling traffic FETCH_DB WRITE_DB --copy-bytes B [--last N]
        HBM bytes per dispatch per kernel from two `rocprofv3 --pmc` runs
        (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each scaled by what
        the same counter reads for k_copy_calibration, a plain copy kernel
        that reads and writes exactly B bytes per launch; JSON
"""
from __future__ import annotations

import argparse
import json
import sqlite3
import sys


def short(name: str) -> str:
    """Kernel symbol without return type, namespace and template arguments."""
    n = name.split("(")[0].strip()
    if n.startswith("void "):
        n = n[5:]
    n = n.replace("dmdx::", "")
    return n.split("<")[0] if not n.startswith("__") else n


def kernel_stats(db: str):
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, total_calls, total_duration, average, percentage from top_kernels "
        "order by total_duration desc").fetchall()
    return [(short(r[0]), int(r[1]), float(r[2]), float(r[3]), float(r[4])) for r in rows]


def window_stats(db: str, steps: int, anchor: str = "k_pair_mixed"):
    """Per-kernel totals over the last `steps` steps: the dispatches that
    start at or after the `steps`-th last dispatch of `anchor` and end by the
    end of its last one.  Returns (wall ns, rows), rows sorted by time."""
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, start, end, stream_id, grid_x, workgroup_x, vgpr_count, lds_size "
        "from kernels order by start").fetchall()
    marks = [r for r in rows if short(r[0]) == anchor]
    if len(marks) < steps:
        raise SystemExit(f"only {len(marks)} dispatches of {anchor}")
    t0, t1 = marks[-steps][1], marks[-1][2]
    agg = {}
    for r in rows:
        if r[1] < t0 or r[2] > t1:
            continue
        a = agg.setdefault(short(r[0]), [0, 0, r[3], r[4], r[5], r[6], r[7]])
        a[0] += 1
        a[1] += r[2] - r[1]
    out = sorted(((k, *v) for k, v in agg.items()), key=lambda x: -x[2])
    return t1 - t0, out


def _counters(last):
    """The counters view, or only each kernel's `last` dispatches (the
    steady state after a warmup) when `last` is set."""
    if not last:
        return "counters_collection", ()
    return ("(select * from (select *, row_number() over (partition by kernel_name, "
            "counter_name order by dispatch_id desc) as rn_ from counters_collection) "
            "where rn_ <= ?)", (int(last),))


def counter_means(db: str, counter: str, last=None):
    con = sqlite3.connect(db)
    src, a = _counters(last)
    rows = con.execute(
        f"select kernel_name, avg(value), count(*) from {src} "
        "where counter_name = ? group by kernel_name", (*a, counter)).fetchall()
    return {short(r[0]): (float(r[1]), int(r[2])) for r in rows}


def all_counter_means(db: str, last=None):
    con = sqlite3.connect(db)
    src, a = _counters(last)
    rows = con.execute(
        f"select kernel_name, counter_name, avg(value) from {src} "
        "group by kernel_name, counter_name", a).fetchall()
    out = {}
    for k, c, v in rows:
        out.setdefault(short(k), {})[c] = float(v)
    return out


def traffic(fetch_db: str, write_db: str, copy_bytes: int, last=None):
    f = counter_means(fetch_db, "FETCH_SIZE", last)
    w = counter_means(write_db, "WRITE_SIZE", last)
    cf = f["k_copy_calibration"][0]
    cw = w["k_copy_calibration"][0]
    sf = copy_bytes / 1024.0 / cf
    sw = copy_bytes / 1024.0 / cw
    out = {"calibration": {"kernel": "k_copy_calibration", "bytes_read": copy_bytes,
                           "bytes_written": copy_bytes, "fetch_kib": cf, "write_kib": cw,
                           "fetch_scale": sf, "write_scale": sw}, "kernels": {}}
    for k in sorted(set(f) | set(w)):
        fb = f.get(k, (0.0, 0))[0] * 1024.0 * sf
        wb = w.get(k, (0.0, 0))[0] * 1024.0 * sw
        out["kernels"][k] = {"fetch_bytes": fb, "write_bytes": wb,
                             "hbm_bytes_per_launch": fb + wb,
                             "dispatches": max(f.get(k, (0, 0))[1], w.get(k, (0, 0))[1])}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("db")
    a.add_argument("--top", type=int, default=25)
    w = sub.add_parser("window")
    w.add_argument("db")
    w.add_argument("--steps", type=int, required=True)
    w.add_argument("--anchor", default="k_pair_mixed")
    c = sub.add_parser("counters")
    c.add_argument("db")
    c.add_argument("--kernel", action="append", default=None)
    c.add_argument("--last", type=int, default=None, help="each kernel's last N dispatches")
    b = sub.add_parser("traffic")
    b.add_argument("fetch_db")
    b.add_argument("write_db")
    b.add_argument("--copy-bytes", type=int, default=1 << 28)
    b.add_argument("--last", type=int, default=None, help="each kernel's last N dispatches")
    args = ap.parse_args(argv)
    if args.cmd == "window":
        wall, rows = window_stats(args.db, args.steps, args.anchor)
        print(f"{args.steps} steps, {wall / 1e6:.2f} ms, {wall / 1e3 / args.steps:.1f} us per step")
        print()
        print("| kernel | calls | us per step | avg us | stream | grid | workgroup | VGPRs | LDS B |")
        print("|---|---:|---:|---:|---:|---:|---:|---:|---:|")
        for k, n, ns, stream, grid, wg, vgpr, lds in rows:
            print(f"| {k} | {n} | {ns / 1e3 / args.steps:.1f} | {ns / 1e3 / n:.1f} | {stream} | "
                  f"{grid} | {wg} | {vgpr} | {lds} |")
        return
    if args.cmd == "stats":
        rows = kernel_stats(args.db)
        # the top_kernels view reports durations in microseconds
        print("| kernel | calls | total ms | avg us | % |")
        print("|---|---:|---:|---:|---:|")
        for name, calls, tot, avg, pct in rows[: args.top]:
            print(f"| {name} | {calls} | {tot / 1e3:.3f} | {avg:.1f} | {pct:.2f} |")
    elif args.cmd == "counters":
        m = all_counter_means(args.db, args.last)
        names = sorted({c for v in m.values() for c in v})
        print("| kernel | " + " | ".join(names) + " |")
        print("|---|" + "---:|" * len(names))
        for k in sorted(m):
            if args.kernel and k not in args.kernel:
                continue
            print(f"| {k} | " + " | ".join(f"{m[k].get(c, float('nan')):.4g}" for c in names) + " |")
    else:
        print(json.dumps(traffic(args.fetch_db, args.write_db, args.copy_bytes, args.last),
                         indent=1))


if __name__ == "__main__":
    sys.exit(main())
