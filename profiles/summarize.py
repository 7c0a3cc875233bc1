"""Summarise a rocprofv3 kernel-trace database (rocpd SQLite) as a markdown table.

Usage: python profiles/summarize.py <run_results.db> "<command that produced it>" > profiles/<name>.md
"""
import sqlite3
import sys


def main(db, cmd):
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    print("# rocprofv3 kernel trace summary\n")
    print("Command (run from /tmp with TMPDIR=/tmp, repository at $R):\n")
    print("```\n%s\n```\n" % cmd)
    print("| kernel | calls | total ms | avg us | % of GPU time |")
    print("|---|---:|---:|---:|---:|")
    for name, n, tot, avg, pct in rows[:15]:  # rocpd top_kernels durations are in microseconds
        short = name.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]
        print("| `%s` | %d | %.2f | %.2f | %.1f |" % (short, n, tot / 1e3, avg, pct))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
