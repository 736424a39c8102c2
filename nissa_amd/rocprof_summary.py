"""Summarise rocprofv3 result databases (rocpd SQLite) as short text tables.

  python -m nissa_amd.rocprof_summary <trace_dir_or_db> [<pmc_dir_or_db> ...] [--match NAME]

For kernel-trace databases: per kernel name, number of dispatches, average and
total duration.  For --pmc databases: per kernel name, the average value per
dispatch of every collected counter (FETCH_SIZE / WRITE_SIZE are in KiB).
"""
from __future__ import annotations

import glob
import os
import sqlite3
import sys
from collections import defaultdict


def _db(path):
    if os.path.isdir(path):
        dbs = glob.glob(os.path.join(path, "**", "*.db"), recursive=True)
        if not dbs:
            raise FileNotFoundError(path)
        return dbs[0]
    return path


def short(name: str) -> str:
    # kernels in an unnamed namespace carry "(anonymous namespace)::" before their
    # name; drop it before cutting the parameter list at the first "("
    n = name.replace("(anonymous namespace)::", "").split("(")[0]
    return n.replace("void ", "").replace("nq::", "")


def kernel_stats(path, match=None):
    c = sqlite3.connect(_db(path))
    rows = c.execute("select name, duration from kernels").fetchall()
    agg = defaultdict(list)
    for n, d in rows:
        if match and match not in n:
            continue
        agg[n].append(d)
    return {n: (len(v), sum(v) / len(v), sum(v)) for n, v in agg.items()}


def pmc_stats(path, match=None):
    c = sqlite3.connect(_db(path))
    rows = c.execute("select name, dispatch_id, counter_name, counter_value from pmc_events").fetchall()
    agg = defaultdict(lambda: defaultdict(list))
    for n, d, cn, v in rows:
        if match and match not in n:
            continue
        agg[n][cn].append(v)
    return {n: {cn: (len(v), sum(v) / len(v)) for cn, v in cs.items()} for n, cs in agg.items()}


def main(argv):
    match = None
    if "--match" in argv:
        i = argv.index("--match")
        match = argv[i + 1]
        argv = argv[:i] + argv[i + 2:]
    for p in argv:
        c = sqlite3.connect(_db(p))
        has_pmc = c.execute("select count(*) from pmc_events").fetchone()[0] > 0
        print(f"# {p}")
        if has_pmc:
            for n, cs in sorted(pmc_stats(p, match).items()):
                vals = ", ".join(f"{cn}={v:.1f} (n={k})" for cn, (k, v) in sorted(cs.items()))
                print(f"{short(n)[:90]:90s} {vals}")
        else:
            ks = kernel_stats(p, match)
            tot = sum(v[2] for v in ks.values()) or 1
            print(f"{'kernel':90s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}")
            for n, (k, avg, s) in sorted(ks.items(), key=lambda kv: -kv[1][2]):
                print(f"{short(n)[:90]:90s} {k:6d} {avg/1e3:10.1f} {s/1e6:10.3f} {100*s/tot:6.1f}")
        print()


if __name__ == "__main__":
    main(sys.argv[1:])
