"""Text summaries of rocprofv3 result databases (SQLite, ``*_results.db``).

    python -m licenta_amd.rocpd_summary DB [--kernels REGEX] [--last N] [--counters]

Without --counters: per kernel, the number of dispatches and the total, average,
minimum and maximum duration in microseconds (the kernel-trace table), sorted by
total time. With --counters: per kernel and counter, the average value per dispatch
(a ``--pmc`` run). --last N restricts each kernel to its last N dispatches, e.g. the
timed repetitions at the end of ``python -m licenta_amd.workload``.
"""
from __future__ import annotations

import argparse
import re
import sqlite3


def _short(name: str) -> str:
    name = name.replace(".kd", "")
    m = re.search(r"(k_\w+)", name)
    return m.group(1) if m else name[:40]


def kernel_table(db: str, pattern: str | None, last: int | None):
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end from kernels order by start").fetchall()
    per: dict[str, list[float]] = {}
    for name, start, end in rows:
        k = _short(name)
        if pattern and not re.search(pattern, k):
            continue
        per.setdefault(k, []).append((end - start) / 1e3)
    out = []
    for k, d in per.items():
        if last:
            d = d[-last:]
        out.append((k, len(d), sum(d), sum(d) / len(d), min(d), max(d)))
    return sorted(out, key=lambda r: -r[2])


def counter_table(db: str, pattern: str | None, last: int | None):
    con = sqlite3.connect(db)
    cols = {r[1] for r in con.execute("pragma table_info(counters_collection)")}
    order = " order by dispatch_id" if "dispatch_id" in cols else ""
    rows = con.execute("select kernel_name, counter_name, value from counters_collection"
                       + order).fetchall()
    per: dict[tuple[str, str], list[float]] = {}
    for kname, cname, value in rows:
        k = _short(kname)
        if pattern and not re.search(pattern, k):
            continue
        per.setdefault((k, cname), []).append(float(value))
    out = []
    for (k, c), v in sorted(per.items()):
        if last:
            v = v[-last:]
        out.append((k, c, len(v), sum(v) / len(v)))
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--kernels", default=None, help="regex on the short kernel name")
    ap.add_argument("--last", type=int, default=None)
    ap.add_argument("--counters", action="store_true")
    args = ap.parse_args()
    if args.counters:
        print(f"{'kernel':22s} {'counter':24s} {'dispatches':>10s} {'avg per dispatch':>18s}")
        for k, c, n, avg in counter_table(args.db, args.kernels, args.last):
            print(f"{k:22s} {c:24s} {n:10d} {avg:18.1f}")
        return
    rows = kernel_table(args.db, args.kernels, args.last)
    total = sum(r[2] for r in rows) or 1.0
    print(f"{'kernel':22s} {'calls':>7s} {'total_us':>11s} {'avg_us':>9s} {'min_us':>9s} "
          f"{'max_us':>9s} {'pct':>6s}")
    for k, n, tot, avg, mn, mx in rows:
        print(f"{k:22s} {n:7d} {tot:11.1f} {avg:9.2f} {mn:9.2f} {mx:9.2f} {100 * tot / total:6.1f}")
    print(f"{'total':22s} {'':7s} {total:11.1f}")


if __name__ == "__main__":
    main()
