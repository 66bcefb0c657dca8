"""Summarise a rocprofv3 kernel trace (rocpd SQLite database or
kernel_stats.csv) into a short markdown table: calls, total and average
duration per kernel, share of device time.

    python -m gpurify_amd.prof_summary <results.db|kernel_stats.csv> "<command>" > out.md
"""
import csv
import sqlite3
import sys


def rows_from(path):
    if path.endswith(".db"):
        con = sqlite3.connect(path)
        # the rocpd top_kernels view reports microseconds
        return [(r[0], int(r[1]), 1e3 * float(r[2]), 1e3 * float(r[3]), float(r[4]))
                for r in con.execute("select name, total_calls, total_duration, average, percentage "
                                     "from top_kernels")]
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]),
                        float(r["AverageNs"]), float(r["Percentage"])))
    return out


def main():
    path, cmd = sys.argv[1], sys.argv[2]
    rows = sorted(rows_from(path), key=lambda r: -r[2])
    total = sum(r[2] for r in rows)
    print(f"Command: `{cmd}`\n")
    print(f"Total kernel time: {total / 1e6:.1f} ms over {sum(r[1] for r in rows)} dispatches\n")
    print("| kernel | calls | total ms | avg us | % |")
    print("|---|---:|---:|---:|---:|")
    for name, calls, tot, avg, pct in rows[:20]:
        short = name.split("(")[0].replace("void ", "")
        print(f"| `{short}` | {calls} | {tot / 1e6:.2f} | {avg / 1e3:.1f} | {pct:.1f} |")


if __name__ == "__main__":
    main()
