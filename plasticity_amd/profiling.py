"""Summaries of rocprofv3 output databases (rocpd SQLite) as short markdown tables.

  python -m plasticity_amd.profiling trace <results.db> "<command>"
      per-kernel calls / total / average duration / share of GPU time (--kernel-trace --stats)
  python -m plasticity_amd.profiling traffic <FETCH_SIZE results.db> <WRITE_SIZE results.db>
      mean FETCH_SIZE and WRITE_SIZE (KiB) per dispatch of every kernel, grouped by grid size;
      the device-to-device copies (__amd_rocclr_copyBuffer) move a known number of bytes and
      calibrate the counters
  python -m plasticity_amd.profiling counters <results.db> <kernel name part> <grid> <n> [<first>]
      every counter of a --pmc run, summed over its per-unit rows, averaged over the last n
      dispatches of the kernel with that grid size, or over the n from the first-th (0-based)
  python -m plasticity_amd.profiling dispatches <results.db> <kernel name part> <n> [<first>]
      durations (us) of the last n dispatches of the kernel (--kernel-trace) and their mean, or of
      the n dispatches from the first-th (0-based, in launch order)
"""
import sqlite3
import sys
from collections import defaultdict


def short_name(name):
    return name.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]


def trace(db, cmd, top=15):
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    print("Command (run from /tmp with TMPDIR=/tmp, repository at $R):\n")
    print("```\n%s\n```\n" % cmd)
    print("| kernel | calls | total ms | avg us | % of GPU time |")
    print("|---|---:|---:|---:|---:|")
    for name, n, tot, avg, pct in rows[:top]:  # rocpd top_kernels durations are in microseconds
        print("| `%s` | %d | %.2f | %.2f | %.1f |" % (short_name(name), n, tot / 1e3, avg, pct))


def _counters(db):
    c = sqlite3.connect(db)
    out = defaultdict(list)
    for name, grid, val in c.execute("select kernel_name, grid_size, value from counters_collection"):
        out[(name, grid)].append(val)
    return out


def traffic(fetch_db, write_db, min_kib=1024):
    f, w = _counters(fetch_db), _counters(write_db)
    print("| kernel | grid | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch |")
    print("|---|---:|---:|---:|---:|")
    for k in sorted(set(f) | set(w), key=lambda k: -sum(f.get(k, [0]))):
        fv, wv = f.get(k, []), w.get(k, [])
        if sum(fv) + sum(wv) < min_kib:
            continue
        print("| `%s` | %d | %d | %.0f | %.0f |" % (short_name(k[0]), k[1], max(len(fv), len(wv)),
                                                   sum(fv) / max(len(fv), 1), sum(wv) / max(len(wv), 1)))


def counters(db, part, grid, n, first=None):
    c = sqlite3.connect(db)
    per = defaultdict(lambda: defaultdict(float))
    for name, g, cname, val, did in c.execute(
            "select kernel_name, grid_size, counter_name, value, dispatch_id from counters_collection"):
        if part in name and g == grid:
            per[did][cname] += val
    last = sorted(per)[-n:] if first is None else sorted(per)[first:first + n]
    which = "the last %d" % len(last) if first is None else "%d from dispatch %d" % (len(last), first)
    print("%d dispatches of `%s` (grid %d); mean over %s:" % (len(per), part, grid, which))
    for cname in sorted({k for d in last for k in per[d]}):
        print("  %s %.6g" % (cname, sum(per[d][cname] for d in last) / len(last)))


def dispatches(db, part, n, first=None):
    c = sqlite3.connect(db)
    rows = [(name, start, end) for name, start, end in
            c.execute("select name, start, end from kernels order by start") if part in name]
    sel = rows[-n:] if first is None else rows[first:first + n]
    durs = [(e - s) / 1e3 for _, s, e in sel]
    which = "the last %d" % len(sel) if first is None else "%d from dispatch %d" % (len(sel), first)
    print("%d dispatches of `%s`; %s: %s us; mean %.2f us" % (len(rows), part, which, ", ".join("%.2f" % d for d in durs),
                                                            sum(durs) / max(len(durs), 1)))


if __name__ == "__main__":
    if len(sys.argv) == 4 and sys.argv[1] == "trace":
        trace(sys.argv[2], sys.argv[3])
    elif len(sys.argv) == 4 and sys.argv[1] == "traffic":
        traffic(sys.argv[2], sys.argv[3])
    elif len(sys.argv) in (5, 6) and sys.argv[1] == "dispatches":
        dispatches(sys.argv[2], sys.argv[3], int(sys.argv[4]), int(sys.argv[5]) if len(sys.argv) == 6 else None)
    elif len(sys.argv) in (6, 7) and sys.argv[1] == "counters":
        counters(sys.argv[2], sys.argv[3], int(sys.argv[4]), int(sys.argv[5]), int(sys.argv[6]) if len(sys.argv) == 7 else None)
    else:
        sys.exit(__doc__)
