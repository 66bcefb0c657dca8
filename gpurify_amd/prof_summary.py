"""Summarise rocprofv3 output into short markdown.

Kernel trace (rocpd SQLite database or kernel_stats.csv): calls, total and
average duration per kernel, share of device time.

    python -m gpurify_amd.prof_summary <results.db|kernel_stats.csv> "<command>" > out.md

Counters (two --pmc runs, FETCH_SIZE then WRITE_SIZE, counter_collection.csv):
per-dispatch means of the kernels matching a symbol, scaled by what the same
counter reads for a streaming copy of known size (each counter reports KiB;
the factor turns the reading into bytes).

    python -m gpurify_amd.prof_summary --pmc <fetch.csv> <write.csv> <copy symbol>
        <copy MiB> <kernel symbol>[=<name match>] <workload> "<command>" > out.md

The last line, `traffic_bytes_per_launch: <symbol> <workload> <bytes>`, is
what bench.py reports as the kernel's HBM traffic.
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


def counter_means(path, counter, symbol):
    """Mean value of `counter` per dispatch of the kernels whose name contains
    `symbol`, and the number of dispatches."""
    vals = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter or symbol not in r["Kernel_Name"]:
                continue
            key = r.get("Dispatch_Id") or r.get("Correlation_Id") or str(len(vals))
            vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"])
    if not vals:
        raise SystemExit(f"no {counter} samples for {symbol} in {path}")
    return sum(vals.values()) / len(vals), len(vals)


def pmc_main(argv):
    fetch_csv, write_csv, copy_sym, copy_mib, sym, workload, cmd = argv
    # "key=match": the traffic line is keyed `key`, the counters are those of
    # the kernels whose name contains `match` (one instantiation of a family)
    key, _, match = sym.partition("=")
    sym = match or key
    copy_bytes = float(copy_mib) * 2 ** 20
    rows = []
    total = 0.0
    for path, counter in ((fetch_csv, "FETCH_SIZE"), (write_csv, "WRITE_SIZE")):
        c_kib, _ = counter_means(path, counter, copy_sym)
        factor = copy_bytes / c_kib          # bytes per counted KiB
        k_kib, nd = counter_means(path, counter, sym)
        b = k_kib * factor
        total += b
        rows.append((counter, c_kib, factor, nd, k_kib, b))
    print(f"Command: `{cmd}`\n")
    print(f"Calibration: `{copy_sym}` streams {copy_mib} MiB in and {copy_mib} MiB out.\n")
    print("| counter | copy reads (KiB) | bytes per KiB | dispatches | kernel mean (KiB) | bytes / launch |")
    print("|---|---:|---:|---:|---:|---:|")
    for counter, c_kib, factor, nd, k_kib, b in rows:
        print(f"| {counter} | {c_kib:.0f} | {factor:.1f} | {nd} | {k_kib:.0f} | {b:.4g} |")
    print(f"\ntraffic_bytes_per_launch: {key} {workload} {total:.6g}")


def main():
    if sys.argv[1] == "--pmc":
        return pmc_main(sys.argv[2:])
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
