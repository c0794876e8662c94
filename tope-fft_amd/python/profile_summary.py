"""Summarise rocprofv3 CSV output into the short text kept under profiles/.

  python tope-fft_amd/python/profile_summary.py stats <kernel_stats.csv>
  python tope-fft_amd/python/profile_summary.py pmc <fetch.csv> <write.csv> \
         <copy_fetch.csv> <copy_write.csv> <copy_bytes>
  python tope-fft_amd/python/profile_summary.py counters <kernel-substring> \
         <counter_collection.csv> [...]

`pmc` scales each kernel's median FETCH_SIZE and WRITE_SIZE (KiB per
dispatch) by the ratio the same counter shows for a plain copy of
`copy_bytes` bytes (which must fetch and write exactly that much), and
prints HBM bytes per launch.
"""
import csv
import statistics
import sys


def short(name):
    return name.split("(")[0].replace("void ", "")


def stats(path):
    rows = list(csv.DictReader(open(path)))
    print("| kernel | calls | avg us | min us | max us | % |")
    print("|---|---|---|---|---|---|")
    for r in rows:
        print(f"| {short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
              f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")


def per_kernel(path):
    d = {}
    for r in csv.DictReader(open(path)):
        d.setdefault(short(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
    return {k: statistics.median(v) for k, v in d.items()}


def pmc(fetch, write, cfetch, cwrite, copy_bytes):
    F, W = per_kernel(fetch), per_kernel(write)
    cf, cw = per_kernel(cfetch), per_kernel(cwrite)
    ck = [k for k in cf if "copy16" in k][0]
    kib = copy_bytes / 1024.0
    ff, fw = kib / cf[ck], kib / cw[ck]
    print(f"copy calibration: {copy_bytes} B copy reads FETCH_SIZE={cf[ck]:.0f} KiB, "
          f"WRITE_SIZE={cw[ck]:.0f} KiB -> scale fetch x{ff:.4f}, write x{fw:.4f}")
    print("| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes/launch (scaled) |")
    print("|---|---|---|---|")
    for k in F:
        b = (F[k] * ff + W.get(k, 0) * fw) * 1024
        print(f"| {k} | {F[k]:.0f} | {W.get(k, 0):.0f} | {b:.0f} |")
    for k in F:
        b = (F[k] * ff + W.get(k, 0) * fw) * 1024
        print(f"traffic_bytes_per_launch {k} {b:.0f}")


def counters(pattern, paths):
    """Median per dispatch of every counter in the given runs, for each
    kernel whose name contains `pattern` (one row per kernel)."""
    vals = {}
    for path in paths:
        for r in csv.DictReader(open(path)):
            k = short(r["Kernel_Name"])
            if pattern in k:
                vals.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    for k, d in vals.items():
        print(f"{k}")
        for c in sorted(d):
            print(f"    {c:32s} {statistics.median(d[c]):16.0f}   (dispatches: {len(d[c])})")


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        stats(sys.argv[2])
    elif sys.argv[1] == "counters":
        counters(sys.argv[2], sys.argv[3:])
    else:
        pmc(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], int(sys.argv[6]))
